// lbf_stages.hip -- the reference's per-stage kernels for gfx950 (one thread
// per node, SoA fields, in-kernel periodic wrap), plus encode/decode,
// reductions and fills used by the C-ABI layer.
//
//   k_norm        compute_norm_interface_kernel (lb_cuda_auxfields_module.f90:19-139)
//   k_div         compute_div_theta_n_kernel    (lb_cuda_auxfields_module.f90:1085-1164)
//   k_phi_bnd     PHI_int_boundary_kernel       (lb_cuda_boundary_module.f90:218-459)
//   k_moments     moments_LB_kernel             (lb_cuda_moments_module.f90:19-341)
//   k_update_phi  update_phifields_kernel       (lb_cuda_update_phi_module.f90:19-104)
//   k_lb_int_bnd  LB_int_boundary_kernel        (lb_cuda_boundary_module.f90:19-216)
// The fused stage is the z-marching kernel of lbf_fused.hip.
//
// These kernels are neighbour-stencils over single fields (at most 27 reads
// of one 4- or 8-byte value per node), HBM-bound; they run once per step on
// the two-component path and are not the dominant cost.
#include "lbf_common.h"
#include "lbf_kernels.h"
#include "lbf_physics.h"

namespace lbf {

static constexpr int BS = 256;

__device__ __forceinline__ bool node_of(size_t t, const DevParams& dp, int& x, int& y, int& z) {
  const size_t n = (size_t)dp.nx * dp.ny * dp.nz;
  if (t >= n) return false;
  x = (int)(t % dp.nx);
  size_t r = t / dp.nx;
  y = (int)(r % dp.ny);
  z = (int)(r / dp.ny);
  return true;
}

#define NB(f, a, b, c) ((R)(f)[nidx(wrapg(x + (a), nx), wrapg(y + (b), ny), wrapg(z + (c), nz), nx, ny)])

// ---------------------------------------------------------------------------
template <int PREC>
__global__ void __launch_bounds__(BS) k_norm(StageArgs a) {
  typedef typename Codec<PREC>::R R;
  typedef typename Aux<PREC>::T TA;
  const DevParams& dp = a.dp;
  int x, y, z;
  if (!node_of((size_t)blockIdx.x * BS + threadIdx.x, dp, x, y, z)) return;
  const int nx = dp.nx, ny = dp.ny, nz = dp.nz;
  const int8_t isf = isf_at(a.isfh, x, y, z, nx, ny, nz);
  if (isf != 1 && isf != -1) return;
  const TA* ph = (const TA*)a.phiS;
  const R p1 = (R)(2.0 / 27.0), p2 = (R)(1.0 / 54.0), p3 = (R)(1.0 / 216.0), p0 = (R)(8.0 / 27.0);
  const R THREE = 3, ONE = 1, TWO = 2;
#define F(a_, b_, c_) NB(ph, a_, b_, c_)
  const R gx = THREE * (p1 * (F(1, 0, 0) - F(-1, 0, 0)) +
                        p2 * ((F(1, 1, 0) - F(-1, -1, 0)) + (F(1, -1, 0) - F(-1, 1, 0)) +
                              (F(1, 0, 1) - F(-1, 0, -1)) + (F(1, 0, -1) - F(-1, 0, 1))) +
                        p3 * ((F(1, 1, 1) - F(-1, -1, -1)) + (F(1, -1, -1) - F(-1, 1, 1)) +
                              (F(1, -1, 1) - F(-1, 1, -1)) + (F(1, 1, -1) - F(-1, -1, 1))));
  const R gy = THREE * (p1 * (F(0, 1, 0) - F(0, -1, 0)) +
                        p2 * ((F(1, 1, 0) - F(-1, -1, 0)) + (F(-1, 1, 0) - F(1, -1, 0)) +
                              (F(0, 1, 1) - F(0, -1, -1)) + (F(0, 1, -1) - F(0, -1, 1))) +
                        p3 * ((F(1, 1, 1) - F(-1, -1, -1)) + (F(-1, 1, -1) - F(1, -1, 1)) +
                              (F(1, 1, -1) - F(-1, -1, 1)) + (F(-1, 1, 1) - F(1, -1, -1))));
  const R gz = THREE * (p1 * (F(0, 0, 1) - F(0, 0, -1)) +
                        p2 * ((F(1, 0, 1) - F(-1, 0, -1)) + (F(-1, 0, 1) - F(1, 0, -1)) +
                              (F(0, 1, 1) - F(0, -1, -1)) + (F(0, -1, 1) - F(0, 1, -1))) +
                        p3 * ((F(1, 1, 1) - F(-1, -1, -1)) + (F(-1, -1, 1) - F(1, 1, -1)) +
                              (F(1, -1, 1) - F(-1, 1, -1)) + (F(-1, 1, 1) - F(1, -1, -1))));
  const R mg = sqrt(gx * gx + gy * gy + gz * gz);
  const R eps = (R)1.0e-9;
  const R c = F(0, 0, 0);
  const size_t n = nidx(x, y, z, nx, ny);
  ((TA*)a.aux[0])[n] = (TA)(gx / (mg + eps));
  ((TA*)a.aux[1])[n] = (TA)(gy / (mg + eps));
  ((TA*)a.aux[2])[n] = (TA)(gz / (mg + eps));
  ((TA*)a.aux[3])[n] = (TA)mg;
  ((TA*)a.aux[4])[n] = (TA)(c * (ONE - c) * (gx / (mg + eps)));
  ((TA*)a.aux[5])[n] = (TA)(c * (ONE - c) * (gy / (mg + eps)));
  ((TA*)a.aux[6])[n] = (TA)(c * (ONE - c) * (gz / (mg + eps)));
  const R lap = (TWO * THREE) *
                (c * (p0 - ONE) +
                 (p1 * (F(1, 0, 0) + F(-1, 0, 0) + F(0, 1, 0) + F(0, -1, 0) + F(0, 0, 1) + F(0, 0, -1)) +
                  p2 * ((F(1, 1, 0) + F(-1, -1, 0)) + (F(1, -1, 0) + F(-1, 1, 0)) + (F(1, 0, 1) + F(-1, 0, -1)) +
                        (F(1, 0, -1) + F(-1, 0, 1)) + (F(0, 1, 1) + F(0, -1, -1)) + (F(0, 1, -1) + F(0, -1, 1))) +
                  p3 * ((F(1, 1, 1) + F(-1, -1, -1)) + (F(1, -1, -1) + F(-1, 1, 1)) + (F(1, -1, 1) + F(-1, 1, -1)) +
                        (F(1, 1, -1) + F(-1, -1, 1)))));
  ((TA*)a.lap)[n] = (TA)lap;
#undef F
}

// ---------------------------------------------------------------------------
template <int PREC>
__global__ void __launch_bounds__(BS) k_div(StageArgs a) {
  typedef typename Codec<PREC>::R R;
  typedef typename Aux<PREC>::T TA;
  const DevParams& dp = a.dp;
  int x, y, z;
  if (!node_of((size_t)blockIdx.x * BS + threadIdx.x, dp, x, y, z)) return;
  const int nx = dp.nx, ny = dp.ny, nz = dp.nz;
  const int8_t isf = isf_at(a.isfh, x, y, z, nx, ny, nz);
  if (isf != 1 && isf != -1) return;
  const TA* ax = (const TA*)a.aux[4];
  const TA* ay = (const TA*)a.aux[5];
  const TA* az = (const TA*)a.aux[6];
  const R p1 = (R)(2.0 / 27.0), p2 = (R)(1.0 / 54.0), p3 = (R)(1.0 / 216.0);
#define X(a_, b_, c_) NB(ax, a_, b_, c_)
#define Y(a_, b_, c_) NB(ay, a_, b_, c_)
#define Z(a_, b_, c_) NB(az, a_, b_, c_)
  const R d = ((p1 * (X(1, 0, 0) - X(-1, 0, 0)) +
                p2 * ((X(1, 1, 0) - X(-1, -1, 0)) + (X(1, -1, 0) - X(-1, 1, 0)) + (X(1, 0, 1) - X(-1, 0, -1)) +
                      (X(1, 0, -1) - X(-1, 0, 1))) +
                p3 * ((X(1, 1, 1) - X(-1, -1, -1)) + (X(1, -1, -1) - X(-1, 1, 1)) + (X(1, -1, 1) - X(-1, 1, -1)) +
                      (X(1, 1, -1) - X(-1, -1, 1)))) +
               (p1 * (Y(0, 1, 0) - Y(0, -1, 0)) +
                p2 * ((Y(1, 1, 0) - Y(-1, -1, 0)) + (Y(-1, 1, 0) - Y(1, -1, 0)) + (Y(0, 1, 1) - Y(0, -1, -1)) +
                      (Y(0, 1, -1) - Y(0, -1, 1))) +
                p3 * ((Y(1, 1, 1) - Y(-1, -1, -1)) + (Y(-1, 1, -1) - Y(1, -1, 1)) + (Y(1, 1, -1) - Y(-1, -1, 1)) +
                      (Y(-1, 1, 1) - Y(1, -1, -1)))) +
               (p1 * (Z(0, 0, 1) - Z(0, 0, -1)) +
                p2 * ((Z(1, 0, 1) - Z(-1, 0, -1)) + (Z(-1, 0, 1) - Z(1, 0, -1)) + (Z(0, 1, 1) - Z(0, -1, -1)) +
                      (Z(0, -1, 1) - Z(0, 1, -1))) +
                p3 * ((Z(1, 1, 1) - Z(-1, -1, -1)) + (Z(-1, -1, 1) - Z(1, 1, -1)) + (Z(1, -1, 1) - Z(-1, 1, -1)) +
                      (Z(-1, 1, 1) - Z(1, -1, -1))))) *
              (R)3;
#undef X
#undef Y
#undef Z
  ((TA*)a.dvt)[nidx(x, y, z, nx, ny)] = (TA)d;
}

// ---------------------------------------------------------------------------
template <int PREC>
__global__ void __launch_bounds__(BS) k_phi_bnd(StageArgs a) {
  typedef typename Codec<PREC>::R R;
  typedef typename Aux<PREC>::T TA;
  const DevParams& dp = a.dp;
  int x, y, z;
  if (!node_of((size_t)blockIdx.x * BS + threadIdx.x, dp, x, y, z)) return;
  const int nx = dp.nx, ny = dp.ny, nz = dp.nz;
  if (isf_at(a.isfh, x, y, z, nx, ny, nz) != 0) return;
  const int nl = dp.lattice == 27 ? 27 : 19;
  const R eps_norm = (R)1.0e-14, huge_cot = (R)1.0e14;
  const R inv_sqrt2 = (R)0.70710678118654752440, inv_sqrt3 = (R)0.57735026918962576450;
  const R half_val = (R)0.5, third_val = (R)(1.0 / 3.0);
  const R pi = (R)3.1415926535897932384626433832795028841971;
  // reference velocity order (vars_module.f90:319-338)
  const int EX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, -1, 1, 1, -1};
  const int EY[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, -1, 1, -1, 1};
  const int EZ[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1};
  TA* ph = (TA*)a.phiS;
  const size_t n = nidx(x, y, z, nx, ny);
  R ghost = (R)ph[n];
  R wnx = 0, wny = 0, wnz = 0;
  for (int l = 1; l < nl; ++l) {
    const int d2 = EX[l] * EX[l] + EY[l] * EY[l] + EZ[l] * EZ[l];
    if (isf_at(a.isfh, x + EX[l], y + EY[l], z + EZ[l], nx, ny, nz) == 0) continue;
    const R s = d2 == 1 ? (R)1 : (d2 == 2 ? half_val : third_val);
    wnx = wnx + s * (-(R)EX[l]);
    wny = wny + s * (-(R)EY[l]);
    wnz = wnz + s * (-(R)EZ[l]);
  }
  const R wn = sqrt(wnx * wnx + wny * wny + wnz * wnz);
  if (wn < eps_norm) return;
  wnx = wnx / wn; wny = wny / wn; wnz = wnz / wn;
  R best = (R)-1.0e30;
  int dx = 0, dy = 0, dz = 0;
  bool have = false;
  R bsx = 0, bsy = 0, bsz = 0;
  for (int l = 1; l < nl; ++l) {
    const int d2 = EX[l] * EX[l] + EY[l] * EY[l] + EZ[l] * EZ[l];
    const int xi = x + EX[l], yj = y + EY[l], zk = z + EZ[l];
    const int8_t v = isf_at(a.isfh, xi, yj, zk, nx, ny, nz);
    if (v != 1 && v != -1) continue;
    const R s = d2 == 1 ? (R)1 : (d2 == 2 ? inv_sqrt2 : inv_sqrt3);
    const R sx = -(R)EX[l], sy = -(R)EY[l], sz = -(R)EZ[l];
    const R score = (sx * wnx + sy * wny + sz * wnz) * s;
    if (score > best) {
      best = score; dx = xi; dy = yj; dz = zk; bsx = sx; bsy = sy; bsz = sz; have = true;
    }
  }
  // the reference skips a donor with a zero (1-based) coordinate, i.e. one in
  // the lower halo (boundary:395); 0-based that is -1
  if (!have || dx == -1 || dy == -1 || dz == -1) return;
  const size_t dn = nidx(wrapg(dx, nx), wrapg(dy, ny), wrapg(dz, nz), nx, ny);
  const R pf = (R)ph[dn];
  const R mg = (R)((const TA*)a.aux[3])[dn];
  const R gx = (R)((const TA*)a.aux[0])[dn] * mg, gy = (R)((const TA*)a.aux[1])[dn] * mg,
          gz = (R)((const TA*)a.aux[2])[dn] * mg;
  const R gdw = gx * wnx + gy * wny + gz * wnz;
  const R tx = gx - gdw * wnx, ty = gy - gdw * wny, tz = gz - gdw * wnz;
  const R tm = sqrt(tx * tx + ty * ty + tz * tz);
  const R th = ((R)180 - (R)dp.wettab) * pi / (R)180;
  const R sn = sin(th), cs = cos(th);
  const R cot = fabs(sn) < eps_norm ? copysign(huge_cot, cs) : cs / sn;
  const R dnn = -tm * cot;
  const R bx = tx + dnn * wnx, by = ty + dnn * wny, bz = tz + dnn * wnz;
  ghost = pf + (bx * bsx + by * bsy + bz * bsz) * ((R)4.0 * ghost * ((R)1.0 - ghost));
  ghost = fmax((R)0, fmin((R)1, ghost));
  ph[n] = (TA)ghost;
}

// ---------------------------------------------------------------------------
template <int PREC>
__global__ void __launch_bounds__(BS) k_moments(StageArgs a) {
  typedef Codec<PREC> C;
  typedef typename C::R R;
  typedef typename C::T T;
  typedef typename Aux<PREC>::T TA;
  const DevParams& dp = a.dp;
  int x, y, z;
  if (!node_of((size_t)blockIdx.x * BS + threadIdx.x, dp, x, y, z)) return;
  const int nx = dp.nx, ny = dp.ny, nz = dp.nz;
  const int8_t isf = isf_at(a.isfh, x, y, z, nx, ny, nz);
  if (isf != 1 && isf != -1) return;
  const size_t n = nidx(x, y, z, nx, ny);
  Mom<R> m = load_mom<PREC>((T* const*)a.S, n, dp);
  const bool two = dp.twocomp != 0;
  R phl = 0, gfx = 0, gfy = 0, gfz = 0, lap = 0, dvt = 0, uox = 0, uoy = 0, uoz = 0;
  if (two) {
    phl = (R)((const TA*)a.phiS)[n];
    const R mg = (R)((const TA*)a.aux[3])[n];
    gfx = (R)((const TA*)a.aux[0])[n] * mg;
    gfy = (R)((const TA*)a.aux[1])[n] * mg;
    gfz = (R)((const TA*)a.aux[2])[n] * mg;
    lap = (R)((const TA*)a.lap)[n];
    dvt = (R)((const TA*)a.dvt)[n];
    if (dp.dens) {
      uox = C::ld((const T*)a.uin[0], n, dp.invscale[1]);
      uoy = C::ld((const T*)a.uin[1], n, dp.invscale[2]);
      uoz = C::ld((const T*)a.uin[2], n, dp.invscale[3]);
    }
  }
  R tFx, tFy, tFz;
  moments_stage<R>(dp, two, phl, gfx, gfy, gfz, lap, dvt, uox, uoy, uoz, m, tFx, tFy, tFz);
  store_mom<PREC>((T* const*)a.S, n, m, dp);
  ((TA*)a.frc[0])[n] = (TA)tFx;
  ((TA*)a.frc[1])[n] = (TA)tFy;
  ((TA*)a.frc[2])[n] = (TA)tFz;
  // the corrected velocity is the next step's u_old (the reference reads it
  // back from this buffer once S and O have swapped)
  const Mom<R> back = load_mom<PREC>((T* const*)a.S, n, dp);
  C::stv((T*)a.uout[0], n, back.u, dp.scale[1]);
  C::stv((T*)a.uout[1], n, back.v, dp.scale[2]);
  C::stv((T*)a.uout[2], n, back.w, dp.scale[3]);
}

// ---------------------------------------------------------------------------
template <int PREC>
__global__ void __launch_bounds__(BS) k_update_phi(StageArgs a) {
  typedef Codec<PREC> C;
  typedef typename C::R R;
  typedef typename C::T T;
  typedef typename Aux<PREC>::T TA;
  const DevParams& dp = a.dp;
  int x, y, z;
  if (!node_of((size_t)blockIdx.x * BS + threadIdx.x, dp, x, y, z)) return;
  const int nx = dp.nx, ny = dp.ny, nz = dp.nz;
  const int8_t isf = isf_at(a.isfh, x, y, z, nx, ny, nz);
  if (isf != 1 && isf != -1) return;
  const size_t n = nidx(x, y, z, nx, ny);
  const R t = -(R)dp.sharp_c * (R)((const TA*)a.dvt)[n];
  const R mg = (R)((const TA*)a.aux[3])[n];
  const R gx = (R)((const TA*)a.aux[0])[n] * mg, gy = (R)((const TA*)a.aux[1])[n] * mg,
          gz = (R)((const TA*)a.aux[2])[n] * mg;
  const R ph = (R)((const TA*)a.phiS)[n];
  const R lap = (R)((const TA*)a.lap)[n];
  const Mom<R> m = load_mom<PREC>((T* const*)a.S, n, dp);
  const R o = ph - m.u * (R)0.5 * (gx) - m.v * (R)0.5 * (gy) - m.w * (R)0.5 * (gz) + (R)dp.tau_diff * lap + t;
  ((TA*)a.phiO)[n] = (TA)o;
}

// ---------------------------------------------------------------------------
template <int PREC>
__global__ void __launch_bounds__(BS) k_lb_int_bnd(StageArgs a) {
  typedef Codec<PREC> C;
  typedef typename C::R R;
  typedef typename C::T T;
  typedef typename Aux<PREC>::T TA;
  const DevParams& dp = a.dp;
  int x, y, z;
  if (!node_of((size_t)blockIdx.x * BS + threadIdx.x, dp, x, y, z)) return;
  const int nx = dp.nx, ny = dp.ny, nz = dp.nz;
  if (isf_at(a.isfh, x, y, z, nx, ny, nz) != -1) return;
  const size_t n = nidx(x, y, z, nx, ny);
  const int EX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, -1, 1, 1, -1};
  const int EY[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, -1, 1, -1, 1};
  const int EZ[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1};
  const int nl = dp.lattice == 27 ? 27 : 19;
  const R cssq = (R)(1.0 / 3.0), invcssq = (R)3, HALF = (R)0.5, ONE = 1, TWO = 2;
  R phl = 0, rhophi = 1;
  if (dp.twocomp) phl = (R)((const TA*)a.phiS)[n];
  if (dp.twocomp && dp.dens) rhophi = (R)dp.rho_r * phl + (ONE - phl) * (R)dp.rho_b;
  Mom<R> s = load_mom<PREC>((T* const*)a.S, n, dp);
  Mom<R> o = load_mom<PREC>((T* const*)a.O, n, dp);
  const R pxx = s.xx - s.u * s.u, pyy = s.yy - s.v * s.v, pzz = s.zz - s.w * s.w;
  const R pxy = s.xy - s.u * s.v, pxz = s.xz - s.u * s.w, pyz = s.yz - s.v * s.w;
  R om;
  if (dp.twocomp) {
    const R visc = ((R)dp.rho_r * (R)dp.visc1 * phl + (ONE - phl) * (R)dp.visc2 * (R)dp.rho_b) / rhophi;
    om = (visc * invcssq + (R)0.5);
    om = (R)1.0 / om;
  } else {
    om = (R)dp.omega;
  }
  for (int l = 1; l < nl; ++l) {
    // upstream node x - c_l (= x + c_opp(l))
    if (isf_at(a.isfh, x - EX[l], y - EY[l], z - EZ[l], nx, ny, nz) != 0) continue;
    const R ex = (R)EX[l], ey = (R)EY[l], ez = (R)EZ[l];
    const int d2 = EX[l] * EX[l] + EY[l] * EY[l] + EZ[l] * EZ[l];
    const R wl = dp.lattice == 27 ? (d2 == 1 ? (R)(2.0 / 27.0) : d2 == 2 ? (R)(1.0 / 54.0) : (R)(1.0 / 216.0))
                                  : (d2 == 1 ? (R)(1.0 / 18.0) : (R)(1.0 / 36.0));
    const R feq = wl * (s.P);
    const R fneq = wl * (HALF / (cssq * cssq)) *
                   ((ex * ex - cssq) * pxx + (ey * ey - cssq) * pyy + (ez * ez - cssq) * pzz + TWO * (ex * ey) * pxy +
                    TWO * (ex * ez) * pxz + TWO * (ey * ez) * pyz);
    const R f = feq + (ONE - om) * fneq;
    o.P = o.P + f;
    o.u = o.u + f * ex; o.v = o.v + f * ey; o.w = o.w + f * ez;
    o.xx = o.xx + f * (ex * ex - cssq);
    o.yy = o.yy + f * (ey * ey - cssq);
    o.zz = o.zz + f * (ez * ez - cssq);
    o.xy = o.xy + f * ex * ey; o.xz = o.xz + f * ex * ez; o.yz = o.yz + f * ey * ez;
  }
  store_mom<PREC>((T* const*)a.O, n, o, dp);
}

// ---------------------------------------------------------------------------
template <int PREC>
static hipError_t stage_t(int stage, const StageArgs& a, hipStream_t s) {
  const size_t n = (size_t)a.dp.nx * a.dp.ny * a.dp.nz;
  const dim3 grid((unsigned)((n + BS - 1) / BS)), block(BS);
  switch (stage) {
    case 1: hipLaunchKernelGGL(k_norm<PREC>, grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_phi_bnd<PREC>, grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL(k_div<PREC>, grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL(k_moments<PREC>, grid, block, 0, s, a); break;
    case 6: hipLaunchKernelGGL(k_update_phi<PREC>, grid, block, 0, s, a); break;
    case 7: hipLaunchKernelGGL(k_lb_int_bnd<PREC>, grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_stage(int stage, const StageArgs& a, int prec, hipStream_t s) {
  switch (prec) {
    case PREC_FP64: return stage_t<PREC_FP64>(stage, a, s);
    case PREC_FP32: return stage_t<PREC_FP32>(stage, a, s);
    case PREC_MIXED: return stage_t<PREC_MIXED>(stage, a, s);
    case PREC_FP16S: return stage_t<PREC_FP16S>(stage, a, s);
    case PREC_BF16S: return stage_t<PREC_BF16S>(stage, a, s);
  }
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// encode / decode
template <int PREC>
__global__ void __launch_bounds__(BS) k_encode(StageArgs a, const float* r0, const float* r1, const float* r2,
                                               const float* r3, const float* r4, const float* r5, const float* r6,
                                               const float* r7, const float* r8, const float* r9) {
  typedef Codec<PREC> C;
  typedef typename C::R R;
  typedef typename C::T T;
  const size_t n = (size_t)blockIdx.x * BS + threadIdx.x;
  if (n >= (size_t)a.dp.nx * a.dp.ny * a.dp.nz) return;
  Mom<R> m;
  m.P = r0[n]; m.u = r1[n]; m.v = r2[n]; m.w = r3[n];
  m.xx = r4[n]; m.yy = r5[n]; m.zz = r6[n]; m.xy = r7[n]; m.xz = r8[n]; m.yz = r9[n];
  store_mom<PREC>((T* const*)a.S, n, m, a.dp);
}

hipError_t launch_encode_moments(void* const* S, const float* const* raw, const DevParams& dp, int prec,
                                 hipStream_t s) {
  StageArgs a{};
  for (int i = 0; i < 10; ++i) a.S[i] = S[i];
  a.dp = dp;
  const size_t n = (size_t)dp.nx * dp.ny * dp.nz;
  const dim3 grid((unsigned)((n + BS - 1) / BS)), block(BS);
  if (prec == PREC_FP16S)
    hipLaunchKernelGGL(k_encode<PREC_FP16S>, grid, block, 0, s, a, raw[0], raw[1], raw[2], raw[3], raw[4], raw[5],
                       raw[6], raw[7], raw[8], raw[9]);
  else if (prec == PREC_BF16S)
    hipLaunchKernelGGL(k_encode<PREC_BF16S>, grid, block, 0, s, a, raw[0], raw[1], raw[2], raw[3], raw[4], raw[5],
                       raw[6], raw[7], raw[8], raw[9]);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

template <int PREC, class O>
__global__ void __launch_bounds__(BS) k_decode(StageArgs a, int comp, O* out) {
  typedef Codec<PREC> C;
  typedef typename C::R R;
  typedef typename C::T T;
  const size_t n = (size_t)blockIdx.x * BS + threadIdx.x;
  if (n >= (size_t)a.dp.nx * a.dp.ny * a.dp.nz) return;
  const Mom<R> m = load_mom<PREC>((T* const*)a.S, n, a.dp);
  const R v[10] = {m.P, m.u, m.v, m.w, m.xx, m.yy, m.zz, m.xy, m.xz, m.yz};
  R r = 0;
#pragma unroll
  for (int i = 0; i < 10; ++i)
    if (i == comp) r = v[i];
  out[n] = (O)r;
}

template <int PREC>
static hipError_t decode_t(const StageArgs& a, int comp, void* out, int dtype, hipStream_t s) {
  const size_t n = (size_t)a.dp.nx * a.dp.ny * a.dp.nz;
  const dim3 grid((unsigned)((n + BS - 1) / BS)), block(BS);
  if (dtype == 0)
    hipLaunchKernelGGL((k_decode<PREC, double>), grid, block, 0, s, a, comp, (double*)out);
  else
    hipLaunchKernelGGL((k_decode<PREC, float>), grid, block, 0, s, a, comp, (float*)out);
  return hipGetLastError();
}

hipError_t launch_decode_field(const void* const* S, int comp, void* out, int out_dtype, const DevParams& dp, int prec,
                               hipStream_t s) {
  StageArgs a{};
  for (int i = 0; i < 10; ++i) a.S[i] = const_cast<void*>(S[i]);
  a.dp = dp;
  switch (prec) {
    case PREC_FP64: return decode_t<PREC_FP64>(a, comp, out, out_dtype, s);
    case PREC_FP32: return decode_t<PREC_FP32>(a, comp, out, out_dtype, s);
    case PREC_MIXED: return decode_t<PREC_MIXED>(a, comp, out, out_dtype, s);
    case PREC_FP16S: return decode_t<PREC_FP16S>(a, comp, out, out_dtype, s);
    case PREC_BF16S: return decode_t<PREC_BF16S>(a, comp, out, out_dtype, s);
  }
  return hipErrorInvalidValue;
}

// elementwise conversion between host dtypes and storage kinds
__device__ __forceinline__ double kload(const void* p, int kind, size_t i, float inv) {
  switch (kind) {
    case K_F64: return ((const double*)p)[i];
    case K_F32: return (double)((const float*)p)[i];
    case K_F16: return (double)(__half2float(((const __half*)p)[i]) * inv);
    default: return (double)(__bfloat162float(((const __hip_bfloat16*)p)[i]) * inv);
  }
}
__device__ __forceinline__ void kstore(void* p, int kind, size_t i, double v, float sc) {
  switch (kind) {
    case K_F64: ((double*)p)[i] = v; break;
    case K_F32: ((float*)p)[i] = (float)v; break;
    case K_F16: ((__half*)p)[i] = __float2half_rn((float)v * sc); break;
    default: ((__hip_bfloat16*)p)[i] = __float2bfloat16((float)v * sc); break;
  }
}
__global__ void __launch_bounds__(BS) k_convert(const void* src, int sk, void* dst, int dk, size_t n, float inv,
                                                float sc) {
  const size_t i = (size_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  kstore(dst, dk, i, kload(src, sk, i, inv), sc);
}

hipError_t launch_convert_in(const void* src, int src_dtype, void* dst, int dst_kind, size_t n, float scale,
                             hipStream_t s) {
  const dim3 grid((unsigned)((n + BS - 1) / BS)), block(BS);
  hipLaunchKernelGGL(k_convert, grid, block, 0, s, src, src_dtype, dst, dst_kind, n, 1.0f, scale);
  return hipGetLastError();
}
hipError_t launch_convert_out(const void* src, int src_kind, void* dst, int dst_dtype, size_t n, float invscale,
                              hipStream_t s) {
  const dim3 grid((unsigned)((n + BS - 1) / BS)), block(BS);
  hipLaunchKernelGGL(k_convert, grid, block, 0, s, src, src_kind, dst, dst_dtype, n, invscale, 1.0f);
  return hipGetLastError();
}

__global__ void __launch_bounds__(BS) k_fill_zero(uint32_t* p, size_t nwords) {
  const size_t i = (size_t)blockIdx.x * BS + threadIdx.x;
  if (i < nwords) p[i] = 0u;
}
hipError_t launch_fill_zero(void* p, size_t bytes, hipStream_t s) {
  const size_t nw = (bytes + 3) / 4;
  if (nw == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((nw + BS - 1) / BS)), dim3(BS), 0, s, (uint32_t*)p, nw);
  return hipGetLastError();
}
hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
  return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s);
}

// ---------------------------------------------------------------------------
// FP64 reductions over |isfluid| == 1 nodes: per-block partials, then one
// block sums the partials in a fixed order (deterministic).
template <int PREC>
__global__ void __launch_bounds__(BS) k_reduce_partial(StageArgs a, double* partial) {
  typedef Codec<PREC> C;
  typedef typename C::R R;
  typedef typename C::T T;
  typedef typename Aux<PREC>::T TA;
  __shared__ double sh[8][BS];
  const DevParams& dp = a.dp;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const size_t N = (size_t)dp.nx * dp.ny * dp.nz;
  for (size_t t = (size_t)blockIdx.x * BS + threadIdx.x; t < N; t += (size_t)gridDim.x * BS) {
    int x, y, z;
    node_of(t, dp, x, y, z);
    const int8_t isf = isf_at(a.isfh, x, y, z, dp.nx, dp.ny, dp.nz);
    if (isf != 1 && isf != -1) continue;
    const Mom<R> m = load_mom<PREC>((T* const*)a.S, t, dp);
    double ph = 0, rho = 1;
    if (dp.twocomp) {
      ph = (double)((const TA*)a.phiS)[t];
      rho = dp.dens ? dp.rho_r * ph + (1.0 - ph) * dp.rho_b : 1.0;
    }
    const double u = (double)m.u, v = (double)m.v, w = (double)m.w;
    acc[0] += (double)m.P;
    acc[1] += rho * u; acc[2] += rho * v; acc[3] += rho * w;
    acc[4] += ph; acc[5] += rho; acc[6] += u * u + v * v + w * w; acc[7] += 1.0;
  }
  for (int k = 0; k < 8; ++k) sh[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int st = BS / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int k = 0; k < 8; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < 8) partial[(size_t)blockIdx.x * 8 + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void k_reduce_final(const double* partial, int nb, double* out) {
  if (threadIdx.x < 8) {
    double s = 0;
    for (int b = 0; b < nb; ++b) s += partial[(size_t)b * 8 + threadIdx.x];
    out[threadIdx.x] = s;
  }
}

hipError_t launch_reduce(const void* const* S, const void* phi, const int8_t* isfh, const DevParams& dp, int prec,
                         double* partial, int nblocks, double* out8, hipStream_t s) {
  StageArgs a{};
  for (int i = 0; i < 10; ++i) a.S[i] = const_cast<void*>(S[i]);
  a.phiS = const_cast<void*>(phi);
  a.isfh = isfh;
  a.dp = dp;
  switch (prec) {
    case PREC_FP64: hipLaunchKernelGGL(k_reduce_partial<PREC_FP64>, dim3(nblocks), dim3(BS), 0, s, a, partial); break;
    case PREC_FP32: hipLaunchKernelGGL(k_reduce_partial<PREC_FP32>, dim3(nblocks), dim3(BS), 0, s, a, partial); break;
    case PREC_MIXED: hipLaunchKernelGGL(k_reduce_partial<PREC_MIXED>, dim3(nblocks), dim3(BS), 0, s, a, partial); break;
    case PREC_FP16S: hipLaunchKernelGGL(k_reduce_partial<PREC_FP16S>, dim3(nblocks), dim3(BS), 0, s, a, partial); break;
    case PREC_BF16S: hipLaunchKernelGGL(k_reduce_partial<PREC_BF16S>, dim3(nblocks), dim3(BS), 0, s, a, partial); break;
    default: return hipErrorInvalidValue;
  }
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(64), 0, s, partial, nblocks, out8);
  return hipGetLastError();
}

}  // namespace lbf
