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
  // isfluid with the reference's 1-node halo; x,y,z unwrapped, clamped to [-1, n]
  x = clampi(x, -1, nx); y = clampi(y, -1, ny); z = clampi(z, -1, nz);
  return isfh[(size_t)(x + 1) + (size_t)(nx + 2) * ((size_t)(y + 1) + (size_t)(ny + 2) * (size_t)(z + 1))];
}

// ---------------------------------------------------------------------------
// Per-node rates shared by the moments stage, the collision and the near-wall
// closure: 1/rho_phi (two-component DENSRATIO; 1 otherwise), the kinematic
// viscosity (the density-weighted two-component mix, or visc1) and the plain
// relaxation rate 1/(3 nu + 1/2).  Computed once per node: the quarter-rate
// reciprocals are not repeated in the branches that use them.
template <class R>
struct NodeRates {
  R irho, visc, om0;
};
template <bool TWO, class R>
__device__ __forceinline__ NodeRates<R> node_rates(const DevParams& dp, R phl) {
  const Phys<R>& ph = dp.ph<R>();
  NodeRates<R> r;
  r.irho = (TWO && dp.dens) ? rcp(ph.rho_r * phl + ((R)1 - phl) * ph.rho_b) : (R)1;
  if (TWO) {
    // without DENSRATIO the reference's viscosity mix still reads rho_r/rho_b
    r.visc = (ph.rho_r * ph.visc1 * phl + ((R)1 - phl) * ph.visc2 * ph.rho_b) * r.irho;
    r.om0 = rcp(r.visc * (R)3 + (R)0.5);
  } else {
    r.visc = ph.visc1;
    r.om0 = ph.omega;
  }
  return r;
}

// ---------------------------------------------------------------------------
// Moments stage of one node (lb_cuda_moments_module.f90:104-339), in place on
// the loaded moments m; returns the stored force tF.  Expression order
// follows the reference; divisions by rho_phi, the INTERFACE_INCOMP
// denominator and tau*cs^2 are multiplications by one reciprocal each, the
// first and last taken from the node's rates (nr = node_rates(dp, phl)).
// rflux: the REPULSIVE_FLUX flux of this node (locaux 6..8), or null.
template <bool TWO, class R>
__device__ __forceinline__ void moments_stage(const DevParams& dp, const NodeRates<R>& nr, R phl, R gfx, R gfy,
                                              R gfz, R lap, R dvt, R uox, R uoy, R uoz, Mom<R>& m, R& tFx, R& tFy,
                                              R& tFz, const R* rflux = nullptr) {
  const Phys<R>& ph = dp.ph<R>();
  const R cssq = (R)(1.0 / 3.0), invcssq = (R)3, HALF = (R)0.5, ONE = (R)1;
  const bool dens = TWO && dp.dens != 0;
  const R rr = ph.rho_r, rb = ph.rho_b;
  R press = m.P;
  R mask = 0, rhophi, irho, Fx, Fy, Fz;
  if (TWO) mask = (R)4.0 * phl * (ONE - phl);
  if (dens) {
    rhophi = rr * phl + (ONE - phl) * rb;
    irho = nr.irho;
    if (dp.buoy) {
      Fx = (rhophi - (rr + rb) * HALF) * ph.fx;
      Fy = (rhophi - (rr + rb) * HALF) * ph.fy;
      Fz = (rhophi - (rr + rb) * HALF) * ph.fz;
    } else {
      Fx = rhophi * ph.fx; Fy = rhophi * ph.fy; Fz = rhophi * ph.fz;
    }
  } else {
    rhophi = (R)1.0; irho = (R)1.0; Fx = ph.fx; Fy = ph.fy; Fz = ph.fz;
  }
  if (TWO) {
    const R st = ((R)4.0 * ph.beta * phl * (phl - (R)1.0) * (phl - (R)0.5) - ph.kapp * lap);
    Fx = Fx + st * gfx; Fy = Fy + st * gfy; Fz = Fz + st * gfz;
    if (rflux) {
      // REPULSIVE_FLUX (moments:148-160): rho_phi * flux clamped to +-1e-3, times rho_phi
      R fr[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        R mt = rflux[c] * rhophi;
        if (fabs(mt) > (R)1.0e-3) mt = (R)1.0e-3 * (mt < (R)0 ? (R)-1 : (R)1);
        fr[c] = mt * rhophi;
      }
      Fx = Fx + fr[0]; Fy = Fy + fr[1]; Fz = Fz + fr[2];
    }
  }
  R grx = 0, gry = 0, grz = 0, vt = 0;
  if (dens) {
    grx = (rr - rb) * gfx; gry = (rr - rb) * gfy; grz = (rr - rb) * gfz;
    Fx = Fx - mask * press * cssq * grx;
    Fy = Fy - mask * press * cssq * gry;
    Fz = Fz - mask * press * cssq * grz;
  }
  tFx = Fx; tFy = Fy; tFz = Fz;
  if (dens) {
    const R pxx = m.xx - uox * uox, pyy = m.yy - uoy * uoy, pzz = m.zz - uoz * uoz;
    const R pxy = m.xy - uox * uoy, pxz = m.xz - uox * uoz, pyz = m.yz - uoy * uoz;
    // visc / (tau cs^2) with tau = 3 visc + 1/2, as 3 visc / tau: 1/tau is
    // the node's plain relaxation rate
    vt = (nr.visc * invcssq) * nr.om0;
    Fx = Fx - mask * vt * (pxx * grx + pxy * gry + pxz * grz);
    Fy = Fy - mask * vt * (pyy * gry + pxy * grx + pyz * grz);
    Fz = Fz - mask * vt * (pzz * grz + pxz * grx + pyz * gry);
  }
  R u = m.u, v = m.v, w = m.w;
  if (dens && dp.incomp) {
    const R t = -ph.sharp_c * dvt;
    u = u + (R)0.5 * Fx * irho;
    v = v + (R)0.5 * Fy * irho;
    w = w + (R)0.5 * Fz * irho;
    const R iden = rcp((R)1.0 - (R)0.5 * (rr - rb) * (ph.tau_diff * lap + t) * irho);
    u = u * iden; v = v * iden; w = w * iden;
    const R a = (rr - rb) * (ph.tau_diff * lap + t);
    tFx = tFx - a * u; tFy = tFy - a * v; tFz = tFz - a * w;
    Fx = Fx - a * u; Fy = Fy - a * v; Fz = Fz - a * w;
  } else {
    u = u + (R)0.5 * Fx * irho;
    v = v + (R)0.5 * Fy * irho;
    w = w + (R)0.5 * Fz * irho;
  }
  m.u = u; m.v = v; m.w = w;
  m.xx = m.xx + Fx * u * irho;
  m.yy = m.yy + Fy * v * irho;
  m.zz = m.zz + Fz * w * irho;
  m.xy = m.xy + HALF * (Fy * u + Fx * v) * irho;
  m.xz = m.xz + HALF * (Fz * u + Fx * w) * irho;
  m.yz = m.yz + HALF * (Fz * v + Fy * w) * irho;
  if (dens) {
    const R pxx = m.xx - u * u, pyy = m.yy - v * v, pzz = m.zz - w * w;
    const R pxy = m.xy - u * v, pxz = m.xz - u * w, pyz = m.yz - v * w;
    tFx = tFx - mask * vt * (pxx * grx + pxy * gry + pxz * grz);
    tFy = tFy - mask * vt * (pyy * gry + pxy * grx + pyz * grz);
    tFz = tFz - mask * vt * (pzz * grz + pxz * grx + pyz * gry);
  }
}

// ---------------------------------------------------------------------------
// Post-collision populations f[cidx(p,q,r)] of one node (the fused kernel's
// collision; also the open-boundary closure's regularised populations).
template <int LAT, bool HO, class R>
__device__ __forceinline__ void collide(const NodeIn<R>& n, R* f) {
  const R om1 = (R)1 - n.om;
  const R u = n.u, v = n.v, w = n.w;
  // order <= 2 Hermite coefficients (post-collision, with half forcing)
  const R A000 = n.P;
  const R A100 = u + (R)0.5 * n.ax, A010 = v + (R)0.5 * n.ay, A001 = w + (R)0.5 * n.az;
  const R A200 = u * u + om1 * n.xx + u * n.ax;
  const R A020 = v * v + om1 * n.yy + v * n.ay;
  const R A002 = w * w + om1 * n.zz + w * n.az;
  const R A110 = u * v + om1 * n.xy + (R)0.5 * (u * n.ay + v * n.ax);
  const R A101 = u * w + om1 * n.xz + (R)0.5 * (u * n.az + w * n.ax);
  const R A011 = v * w + om1 * n.yz + (R)0.5 * (v * n.az + w * n.ay);
  if constexpr (LAT == 27) {
    R A[27];  // A[a + 3b + 9c]
#define AI(a, b, c) A[(a) + 3 * (b) + 9 * (c)]
    AI(0, 0, 0) = A000;
    AI(1, 0, 0) = A100; AI(0, 1, 0) = A010; AI(0, 0, 1) = A001;
    AI(2, 0, 0) = A200; AI(0, 2, 0) = A020; AI(0, 0, 2) = A002;
    AI(1, 1, 0) = A110; AI(1, 0, 1) = A101; AI(0, 1, 1) = A011;
    if constexpr (HO) {
      const R pxx = n.xx, pyy = n.yy, pzz = n.zz, pxy = n.xy, pxz = n.xz, pyz = n.yz;
      const R uu = u * u, vv = v * v, ww = w * w;
      // E_k and N_k of fused:668-702 (k = 10..26), N_k transcribed as written
      AI(2, 1, 0) = uu * v + om1 * ((R)2 * pxy * u + pxx * v);
      AI(2, 0, 1) = uu * w + om1 * ((R)2 * pxz * u + pxx * w);
      AI(1, 2, 0) = u * vv + om1 * (pyy * u + (R)2 * pxy * v);
      AI(1, 0, 2) = u * ww + om1 * (pzz * u + (R)2 * pxz * w);
      AI(0, 1, 2) = v * ww + om1 * (pzz * v + (R)2 * pyz * w);
      AI(0, 2, 1) = vv * w + om1 * ((R)2 * pyz * v + pyy * w);
      AI(1, 1, 1) = u * v * w + om1 * (pyz * u + pxz * v + pxy * w);
      AI(2, 2, 0) = uu * vv + om1 * (pyy * uu + (R)2 * v * ((R)3 * pxy * u + pxx * v));
      AI(2, 0, 2) = uu * ww + om1 * (pzz * uu + (R)2 * w * ((R)3 * pxz * u + pxx * w));
      AI(0, 2, 2) = vv * ww + om1 * (pzz * vv + (R)2 * w * ((R)3 * pyz * v + pyy * w));
      AI(1, 1, 2) = u * v * ww + om1 * (pzz * u * v + w * ((R)3 * pyz * u + (R)3 * pxz * v + (R)2 * pxy * w));
      AI(1, 2, 1) = u * vv * w + om1 * ((R)2 * pyz * u * v + pxz * vv + (R)2 * pyy * u * w + (R)4 * pxy * v * w);
      AI(2, 1, 1) = uu * v * w + om1 * (pyz * uu + (R)2 * pxz * u * v + (R)4 * pxy * u * w + (R)2 * pxx * v * w);
      AI(2, 1, 2) = uu * v * ww +
                    om1 * (pzz * uu * v + (R)3 * w * (pyz * uu + (R)2 * pxz * u * v + (R)2 * pxy * u * w + pxx * v * w));
      AI(2, 2, 1) = uu * vv * w +
                    om1 * ((R)2 * u * v * (pyz * u + pxz * v) + (R)2 * pyy * uu * w + v * ((R)10 * pxy * u + (R)3 * pxx * v) * w);
      AI(1, 2, 2) = u * vv * ww +
                    om1 * (pzz * u * vv + (R)3 * w * ((R)2 * pyz * u * v + pxz * vv + pyy * u * w + (R)2 * pxy * v * w));
      AI(2, 2, 2) = uu * vv * ww +
                    om1 * (pzz * uu * vv + w * ((R)6 * u * v * (pyz * u + pxz * v) + (R)3 * pyy * uu * w +
                                                (R)2 * v * ((R)7 * pxy * u + (R)2 * pxx * v) * w));
    } else {
      AI(2, 1, 0) = 0; AI(2, 0, 1) = 0; AI(1, 2, 0) = 0; AI(1, 0, 2) = 0; AI(0, 1, 2) = 0; AI(0, 2, 1) = 0;
      AI(1, 1, 1) = 0; AI(2, 2, 0) = 0; AI(2, 0, 2) = 0; AI(0, 2, 2) = 0; AI(1, 1, 2) = 0; AI(1, 2, 1) = 0;
      AI(2, 1, 1) = 0; AI(2, 1, 2) = 0; AI(2, 2, 1) = 0; AI(1, 2, 2) = 0; AI(2, 2, 2) = 0;
    }
#undef AI
    // x pass: A[a][b][c] -> B[p][b][c]
    R B[27];
#pragma unroll
    for (int bc = 0; bc < 9; ++bc) recon3(A[3 * bc + 0], A[3 * bc + 1], A[3 * bc + 2], B[3 * bc + 0], B[3 * bc + 1], B[3 * bc + 2]);
    // y pass: B[p][b][c] -> A[p][q][c]
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int pp = 0; pp < 3; ++pp)
        recon3(B[pp + 0 + 9 * c], B[pp + 3 + 9 * c], B[pp + 6 + 9 * c], A[pp + 0 + 9 * c], A[pp + 3 + 9 * c], A[pp + 6 + 9 * c]);
    // z pass: A[p][q][c] -> f[p][q][r]
#pragma unroll
    for (int pq = 0; pq < 9; ++pq) recon3(A[pq], A[pq + 9], A[pq + 18], f[pq], f[pq + 9], f[pq + 18]);
  } else {
    // D3Q19: f_l = w_l [A0 + 3 c.A1 + 4.5 sum_ab (c_a c_b - delta/3) A2_ab]
#pragma unroll
    for (int r = -1; r <= 1; ++r)
#pragma unroll
      for (int q = -1; q <= 1; ++q)
#pragma unroll
        for (int p = -1; p <= 1; ++p) {
          const int c = cidx(p, q, r);
          if (!present<LAT>(p, q, r)) { f[c] = 0; continue; }
          const R wl = (R)weight<LAT>(p, q, r);
          R s = A000;
          if (p) s += (R)(3 * p) * A100;
          if (q) s += (R)(3 * q) * A010;
          if (r) s += (R)(3 * r) * A001;
          s += (R)(4.5 * (p * p - 1.0 / 3.0)) * A200;
          s += (R)(4.5 * (q * q - 1.0 / 3.0)) * A020;
          s += (R)(4.5 * (r * r - 1.0 / 3.0)) * A002;
          if (p * q) s += (R)(9 * p * q) * A110;
          if (p * r) s += (R)(9 * p * r) * A101;
          if (q * r) s += (R)(9 * q * r) * A011;
          f[c] = wl * s;
        }
  }
}

// relaxation rate of the fused / LB_int_boundary kernels: omega, or
// 1/(3 nu(phi) + 1/2) with nu(phi) the density-weighted viscosity mix;
// under SMAGORINSKI (fused:198-210, 603-616; boundary:145-160)
//   tau = 1/2 + (3 nu + sqrt((3 nu)^2 + C sqrt(2 Q) / rho_phi)) / 6,
// Q = Pneq:Pneq, C = 0.053*18.0 evaluated in single precision as the
// reference's default-real literals are.
template <bool TWO, class R>
__device__ __forceinline__ R relax_rate(const DevParams& dp, const NodeRates<R>& nr, R pxx, R pyy, R pzz, R pxy,
                                        R pxz, R pyz) {
  if (dp.smag) {
    const R QQ = pxx * pxx + pyy * pyy + pzz * pzz + (R)2 * (pxy * pxy + pxz * pxz + pyz * pyz);
    const R C = (R)(0.053f * 18.0f);
    const R t3 = (R)3 * nr.visc;
    return rcp((R)0.5 + (R)(1.0 / 6.0) * ((R)3 * nr.visc + sqrt(t3 * t3 + C * sqrt((R)2 * QQ) * nr.irho)));
  }
  return nr.om0;
}

}  // namespace lbf
