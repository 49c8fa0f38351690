// lbf_physics.h -- per-node physics shared by the fused and stage kernels.
#pragma once
#include "lbf_common.h"

namespace lbf {

template <class R>
struct NodeIn {  // collision input of one node
  R P, u, v, w, xx, yy, zz, xy, xz, yz;  // P, velocity, non-equilibrium Pi - u u
  R ax, ay, az;                          // force / rho_phi
  R om;                                  // relaxation rate
};

__device__ __forceinline__ int wrapg(int i, int n) {
  // periodic wrap for i in [-n, 2n)
  i = i < 0 ? i + n : i;
  return i >= n ? i - n : i;
}
__device__ __forceinline__ int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }

__device__ __forceinline__ int8_t isf_at(const int8_t* isfh, int x, int y, int z, int nx, int ny, int nz) {
  // isfluid with the reference's 1-node halo; x,y,z unwrapped in [-1, n]
  x = clampi(x, -1, nx); y = clampi(y, -1, ny); z = clampi(z, -1, nz);
  return isfh[(size_t)(x + 1) + (size_t)(nx + 2) * ((size_t)(y + 1) + (size_t)(ny + 2) * (size_t)(z + 1))];
}

// ---------------------------------------------------------------------------
// Moments stage of one node (lb_cuda_moments_module.f90:104-339), in place on
// the loaded moments m; returns the stored force tF.  Expression order
// follows the reference.
template <class R>
__device__ __forceinline__ void moments_stage(const DevParams& dp, bool two, R phl, R gfx, R gfy, R gfz,
                                              R lap, R dvt, R uox, R uoy, R uoz, Mom<R>& m, R& tFx, R& tFy,
                                              R& tFz) {
  const R cssq = (R)(1.0 / 3.0), invcssq = (R)3, HALF = (R)0.5, ONE = (R)1;
  const R rr = (R)dp.rho_r, rb = (R)dp.rho_b;
  const bool dens = dp.dens != 0;
  R press = m.P;
  R mask = 0, rhophi, Fx, Fy, Fz;
  if (two) mask = (R)4.0 * phl * (ONE - phl);
  if (dens) {
    rhophi = rr * phl + (ONE - phl) * rb;
    if (dp.buoy) {
      Fx = (rhophi - (rr + rb) * HALF) * (R)dp.fx;
      Fy = (rhophi - (rr + rb) * HALF) * (R)dp.fy;
      Fz = (rhophi - (rr + rb) * HALF) * (R)dp.fz;
    } else {
      Fx = rhophi * (R)dp.fx; Fy = rhophi * (R)dp.fy; Fz = rhophi * (R)dp.fz;
    }
  } else {
    rhophi = (R)1.0; Fx = (R)dp.fx; Fy = (R)dp.fy; Fz = (R)dp.fz;
  }
  if (two) {
    R st = ((R)4.0 * (R)dp.beta * phl * (phl - (R)1.0) * (phl - (R)0.5) - (R)dp.kapp * lap);
    Fx = Fx + st * gfx; Fy = Fy + st * gfy; Fz = Fz + st * gfz;
  }
  R grx = 0, gry = 0, grz = 0;
  if (dens) {
    grx = (rr - rb) * gfx; gry = (rr - rb) * gfy; grz = (rr - rb) * gfz;
    Fx = Fx - mask * press * cssq * grx;
    Fy = Fy - mask * press * cssq * gry;
    Fz = Fz - mask * press * cssq * grz;
  }
  tFx = Fx; tFy = Fy; tFz = Fz;
  R visc = 0, tau = 0;
  if (dens) {
    R pxx = m.xx - uox * uox, pyy = m.yy - uoy * uoy, pzz = m.zz - uoz * uoz;
    R pxy = m.xy - uox * uoy, pxz = m.xz - uox * uoz, pyz = m.yz - uoy * uoz;
    visc = (rr * (R)dp.visc1 * phl + ((R)1.0 - phl) * (R)dp.visc2 * rb) / rhophi;
    tau = (visc * invcssq + HALF);
    Fx = Fx - mask * (visc / (tau * cssq)) * (pxx * grx + pxy * gry + pxz * grz);
    Fy = Fy - mask * (visc / (tau * cssq)) * (pyy * gry + pxy * grx + pyz * grz);
    Fz = Fz - mask * (visc / (tau * cssq)) * (pzz * grz + pxz * grx + pyz * gry);
  }
  R u = m.u, v = m.v, w = m.w;
  if (dens && dp.incomp) {
    R t = -(R)dp.sharp_c * dvt;
    u = u + (R)0.5 * Fx / (rhophi);
    v = v + (R)0.5 * Fy / (rhophi);
    w = w + (R)0.5 * Fz / (rhophi);
    R den = ((R)1.0 - (R)0.5 * (rr - rb) * ((R)dp.tau_diff * lap + t) / rhophi);
    u = u / den; v = v / den; w = w / den;
    R a = (rr - rb) * ((R)dp.tau_diff * lap + t);
    tFx = tFx - a * u; tFy = tFy - a * v; tFz = tFz - a * w;
    Fx = Fx - a * u; Fy = Fy - a * v; Fz = Fz - a * w;
  } else {
    u = u + (R)0.5 * Fx / rhophi;
    v = v + (R)0.5 * Fy / rhophi;
    w = w + (R)0.5 * Fz / rhophi;
  }
  m.u = u; m.v = v; m.w = w;
  m.xx = m.xx + Fx * u / rhophi;
  m.yy = m.yy + Fy * v / rhophi;
  m.zz = m.zz + Fz * w / rhophi;
  m.xy = m.xy + HALF * (Fy * u + Fx * v) / rhophi;
  m.xz = m.xz + HALF * (Fz * u + Fx * w) / rhophi;
  m.yz = m.yz + HALF * (Fz * v + Fy * w) / rhophi;
  if (dens) {
    R pxx = m.xx - u * u, pyy = m.yy - v * v, pzz = m.zz - w * w;
    R pxy = m.xy - u * v, pxz = m.xz - u * w, pyz = m.yz - v * w;
    tFx = tFx - mask * (visc / (tau * cssq)) * (pxx * grx + pxy * gry + pxz * grz);
    tFy = tFy - mask * (visc / (tau * cssq)) * (pyy * gry + pxy * grx + pyz * grz);
    tFz = tFz - mask * (visc / (tau * cssq)) * (pzz * grz + pxz * grx + pyz * gry);
  }
}

}  // namespace lbf
