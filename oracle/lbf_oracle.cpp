// lbf_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// A CPU restatement of the reference's moment-represented lattice-Boltzmann
// time step, used by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg as the checker and the timed CPU baseline.  The product
// (liblbfast_hip.so) never links, loads or calls this file.
//
// It follows the reference stage by stage on a plain (i,j,k) grid with the
// reference's 1-node halo, halo-copy semantics, ping-pong buffers, in-place
// moments update and arithmetic order:
//   lattice tables          vars_module.f90:316-338 (D3Q27), 666-684 (D3Q19),
//                           340-664 (Minv_mrt / M_mrt), 708-711 (D3Q27 weights)
//   derived parameters      LBFAST.f90:212-219
//   isfluid classification  LBFAST.f90:258-328
//   compute_norm            lb_cuda_auxfields_module.f90:19-139
//   div_theta_n             lb_cuda_auxfields_module.f90:1085-1164
//   PHI_int_boundary        lb_cuda_boundary_module.f90:218-459
//   moments                 lb_cuda_moments_module.f90:19-341
//   fused stream+collide    lb_cuda_fused_module.f90:426-750
//   update_phi              lb_cuda_update_phi_module.f90:19-104
//   LB_int_boundary         lb_cuda_boundary_module.f90:19-216
//   BOUNCE_BACK fused       lb_cuda_fused_module.f90:20-424 (fused_LB_kernel_bb)
//   SMAGORINSKI relaxation  lb_cuda_fused_module.f90:198-210, 603-616;
//                           lb_cuda_boundary_module.f90:145-160
//   REPULSIVE_FLUX          lb_cuda_repulsive_module.f90:19-292 and
//                           lb_cuda_moments_module.f90:148-160
//   halo (intpbc) copies    mpi_module.f90:646-709, 2820-2918
//   stage order             integrator_module.f90:179-233, 266-545
//   initial O := S          init_conditions_module.f90:584-587
//
// Templated on compute precision R (the reference's `db`) and storage
// precision T (`strdb`): <double,double> = PRC 8, <float,float> = PRC 4,
// <double,float> = MIXEDPRC with STRPRC 4.
//
// Parity pins (checked by tests/test_oracle_*.py): Minv_mrt against the
// reference's literal table and M*Minv = I; rest-state and uniform-flow
// fixed points; low-order moment closure of the rebuilt populations;
// streaming as an exact permutation; mass conservation and exact momentum
// injection; isotropic stencils exact on linear/quadratic phi; Taylor-Green
// kinetic-energy decay; Laplace pressure jump.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <string>
#include <algorithm>
#include <chrono>
#include <limits>
#include "../include/lbfast_hip.h"

namespace orc {

// ---------------------------------------------------------------- lattice
// D3Q27 velocity ordering of vars_module.f90:319-338; the first 19 entries
// are the D3Q19 set of vars_module.f90:666-684.
static const int EX[27] = {0, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1, 0, 0, 0, 0, 1,-1,-1, 1, 1,-1, 1,-1,-1, 1, 1,-1};
static const int EY[27] = {0, 0, 0, 1,-1, 0, 0, 1,-1,-1, 1, 1,-1, 1,-1, 0, 0, 0, 0, 1,-1,-1, 1,-1, 1,-1, 1};
static const int EZ[27] = {0, 0, 0, 0, 0, 1,-1, 0, 0, 0, 0, 1,-1,-1, 1, 1,-1, 1,-1, 1,-1, 1,-1, 1,-1,-1, 1};
static const int OPP[27] = {0, 2, 1, 4, 3, 6, 5, 8, 7,10, 9,12,11,14,13,16,15,18,17,20,19,22,21,24,23,26,25};

// Hermite basis index k -> per-axis order (a,b,c), from the rows of M_mrt
// (vars_module.f90:522-664): 0:1 1:x 2:y 3:z 4:xx 5:yy 6:zz 7:xy 8:xz 9:yz
// 10:xxy 11:xxz 12:xyy 13:xzz 14:yzz 15:yyz 16:xyz 17:xxyy 18:xxzz 19:yyzz
// 20:xyzz 21:xyyz 22:xxyz 23:xxyzz 24:xxyyz 25:xyyzz 26:xxyyzz, where an
// order-2 factor is (c^2 - 1/3).
static const int HK[27][3] = {
  {0,0,0},{1,0,0},{0,1,0},{0,0,1},{2,0,0},{0,2,0},{0,0,2},{1,1,0},{1,0,1},{0,1,1},
  {2,1,0},{2,0,1},{1,2,0},{1,0,2},{0,1,2},{0,2,1},{1,1,1},{2,2,0},{2,0,2},{0,2,2},
  {1,1,2},{1,2,1},{2,1,1},{2,1,2},{2,2,1},{1,2,2},{2,2,2}};

static double herm1(int order, int c) {
  if (order == 0) return 1.0;
  if (order == 1) return (double)c;
  return (double)(c * c) - 1.0 / 3.0;
}
static double hnorm1(int order) {  // 1, cs^2, 2 cs^4
  if (order == 0) return 1.0;
  if (order == 1) return 1.0 / 3.0;
  return 2.0 / 9.0;
}
static double w1d(int c) { return c == 0 ? 2.0 / 3.0 : 1.0 / 6.0; }

// Minv_mrt(k,l) = w_l H_k(c_l) / b_k, the closed form of the reference's
// literal table (vars_module.f90:359-521); tests compare the two entry by entry.
double minv_d3q27(int k, int l) {
  const int c[3] = {EX[l], EY[l], EZ[l]};
  double v = 1.0;
  for (int d = 0; d < 3; ++d) v *= w1d(c[d]) * herm1(HK[k][d], c[d]) / hnorm1(HK[k][d]);
  return v;
}
double m_d3q27(int k, int l) {  // M_mrt: H_k(c_l)
  const int c[3] = {EX[l], EY[l], EZ[l]};
  double v = 1.0;
  for (int d = 0; d < 3; ++d) v *= herm1(HK[k][d], c[d]);
  return v;
}

struct Lat {
  int nl;        // nlinks + 1
  double w[27];  // p(l)
};
static Lat make_lat(int lattice) {
  Lat L{};
  if (lattice == 27) {
    L.nl = 27;
    for (int l = 0; l < 27; ++l) {
      int n2 = EX[l] * EX[l] + EY[l] * EY[l] + EZ[l] * EZ[l];
      L.w[l] = n2 == 0 ? 8.0 / 27.0 : n2 == 1 ? 2.0 / 27.0 : n2 == 2 ? 1.0 / 54.0 : 1.0 / 216.0;
    }
  } else {
    L.nl = 19;
    for (int l = 0; l < 19; ++l) {
      int n2 = EX[l] * EX[l] + EY[l] * EY[l] + EZ[l] * EZ[l];
      L.w[l] = n2 == 0 ? 1.0 / 3.0 : n2 == 1 ? 1.0 / 18.0 : 1.0 / 36.0;
    }
  }
  return L;
}

// ---------------------------------------------------------------- params
// LBFAST.f90:212-219
static void derive(lbf_params* p) {
  p->tau1 = p->visc1 * 3.0 + 0.5;
  p->omega = 1.0 / p->tau1;
  if (p->twocomponent) {
    p->beta = 12.0 * p->sigma / p->width;
    p->kapp = 1.5 * p->sigma * p->width;
    p->sharp_c = 4.0 * p->tau_diff / p->width;
    p->tau2 = p->visc2 * 3.0 + 0.5;
  }
}

// LBFAST.f90:258-328 on an interior array (Fortran order) -> classified interior.
static int classify(const lbf_params* p, const int8_t* in, int8_t* out) {
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  const int hx = nx + 2, hy = ny + 2, hz = nz + 2;
  std::vector<int8_t> a((size_t)hx * hy * hz, 1);  // isfluid = 1, halo included
  auto H = [&](int i, int j, int k) -> int8_t& { return a[(size_t)i + (size_t)hx * (j + (size_t)hy * k)]; };
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        int8_t v = in ? in[(size_t)(i - 1) + (size_t)nx * ((j - 1) + (size_t)ny * (k - 1))] : 1;
        H(i, j, k) = v;
      }
  if (p->internal_obstacles) {
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
          if (!p->pbc[0] && (i == 1 || i == nx)) H(i, j, k) = 0;
          if (!p->pbc[1] && (j == 1 || j == ny)) H(i, j, k) = 0;
          if (!p->pbc[2] && (k == 1 || k == nz)) H(i, j, k) = 0;
        }
  }
  // exchange_isf_intpbc: periodic halo copy; non-periodic halo stays 1
  auto wrap = [](int i, int n) { return i < 1 ? i + n : (i > n ? i - n : i); };
  for (int k = 0; k < hz; ++k)
    for (int j = 0; j < hy; ++j)
      for (int i = 0; i < hx; ++i) {
        bool hal = i == 0 || i == hx - 1 || j == 0 || j == hy - 1 || k == 0 || k == hz - 1;
        if (!hal) continue;
        bool ok = true;
        if ((i == 0 || i == hx - 1) && !p->pbc[0]) ok = false;
        if ((j == 0 || j == hy - 1) && !p->pbc[1]) ok = false;
        if ((k == 0 || k == hz - 1) && !p->pbc[2]) ok = false;
        if (ok) H(i, j, k) = H(wrap(i, nx), wrap(j, ny), wrap(k, nz));
      }
  const int nl = p->lattice == 27 ? 27 : 19;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        if (H(i, j, k) == 1) {
          for (int l = 1; l < nl; ++l) {
            int8_t v = H(i + EX[l], j + EY[l], k + EZ[l]);
            if (v != 1 && v != -1) H(i, j, k) = -1;
          }
        }
      }
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i)
        out[(size_t)(i - 1) + (size_t)nx * ((j - 1) + (size_t)ny * (k - 1))] = H(i, j, k);
  return 0;
}

// ---------------------------------------------------------------- simulator
struct SimBase {
  virtual ~SimBase() {}
  virtual int set_field(int id, const double* src) = 0;
  virtual int get_field(int id, double* dst) = 0;
  virtual int set_isfluid(const int8_t* src) = 0;
  virtual int init_state() = 0;
  virtual int begin_step() = 0;
  virtual int stage(int id) = 0;
  virtual int end_step() = 0;
  virtual int step(int n) = 0;
  virtual int populations(double* dst) = 0;
  std::string err;
};

template <class R, class T>
struct Sim : SimBase {
  lbf_params p;
  Lat L;
  int nx, ny, nz, hx, hy, hz;
  size_t N;
  std::vector<T> h[2][10];
  std::vector<T> phi[2];
  std::vector<T> aux[7], loc[8], frc[3];  // loc = locaux 1..8
  std::vector<int8_t> isf, rep_mask;
  std::vector<R> pop;  // post-collision populations of the last fused stage
  int cur = 0;         // index of the S buffer
  int stepno = 0;
  bool two, dens, incomp, obst, ho, bb, smag, rep, lagr, obc;
  double phi_sum_ini = 0;  // BCPHIFLUX: global_phi_sum_ini
  R Minv[27][27];

  explicit Sim(const lbf_params& pp) : p(pp) {
    derive(&p);
    L = make_lat(p.lattice);
    nx = p.nx; ny = p.ny; nz = p.nz;
    hx = nx + 2; hy = ny + 2; hz = nz + 2;
    N = (size_t)hx * hy * hz;
    for (int b = 0; b < 2; ++b) {
      for (int f = 0; f < 10; ++f) h[b][f].assign(N, T(0));
      phi[b].assign(N, T(0));
    }
    for (auto& a : aux) a.assign(N, T(0));
    for (auto& a : loc) a.assign(N, T(0));
    for (auto& a : frc) a.assign(N, T(0));
    isf.assign(N, 1);
    rep_mask.assign(N, 0);
    two = p.twocomponent != 0;
    dens = two && p.densratio != 0;
    incomp = dens && p.interface_incomp != 0;
    obst = p.internal_obstacles != 0;
    ho = p.highorder != 0 && p.lattice == 27;
    bb = obst && p.bounce_back != 0;
    smag = p.smagorinsky != 0;
    rep = two && p.repulsive_flux != 0;
    lagr = two && obst && p.bcphiflux != 0;
    obc = obst && p.openbc != 0;
    for (int k = 0; k < 27; ++k)
      for (int l = 0; l < 27; ++l) Minv[k][l] = (R)minv_d3q27(k, l);
  }

  inline size_t I(int i, int j, int k) const { return (size_t)i + (size_t)hx * (j + (size_t)hy * k); }
  static inline R ld(const std::vector<T>& a, size_t n) { return (R)a[n]; }
  static inline T st(R v) { return (T)v; }

  // SMAGORINSKI relaxation rate (fused:603-616): tau = 1/2 + (3 nu +
  // sqrt((3 nu)^2 + C sqrt(2 Q) / rho_phi)) / 6 with Q = Pneq:Pneq.  The
  // reference writes C as 0.053*18.0, a product of two default (single
  // precision) reals, so C is that float32 product widened to double.
  static R smag_omega(R visc, R pxx, R pyy, R pzz, R pxy, R pxz, R pyz, R inv_rho) {
    const R QQ = pxx * pxx + pyy * pyy + pzz * pzz + (R)2 * (pxy * pxy + pxz * pxz + pyz * pyz);
    const R C = (R)(double)(0.053f * 18.0f);
    const R t3 = (R)3.0 * visc;
    R tau = (R)0.5 + ((R)1.0 / (R)6.0) * ((R)3.0 * visc + std::sqrt(t3 * t3 + C * std::sqrt((R)2.0 * QQ) * inv_rho));
    return (R)1.0 / tau;
  }

  // ---- halo (intpbc) ----------------------------------------------------
  // A halo node is copied from its periodic image only when every axis on
  // which it lies outside [1,n] is periodic (mpi_module.f90:646-709).
  void halo(std::vector<T>& a) {
    auto wrap = [](int i, int n) { return i < 1 ? i + n : (i > n ? i - n : i); };
    for (int k = 0; k < hz; ++k)
      for (int j = 0; j < hy; ++j)
        for (int i = 0; i < hx; ++i) {
          bool ox = i == 0 || i == hx - 1, oy = j == 0 || j == hy - 1, oz = k == 0 || k == hz - 1;
          if (!(ox || oy || oz)) continue;
          if ((ox && !p.pbc[0]) || (oy && !p.pbc[1]) || (oz && !p.pbc[2])) continue;
          a[I(i, j, k)] = a[I(wrap(i, nx), wrap(j, ny), wrap(k, nz))];
        }
  }

  // ---- compute_norm (lb_cuda_auxfields_module.f90:62-135) -----------------
  void compute_norm(const std::vector<T>& ph) {
    const R p1 = (R)(2.0 / 27.0), p2 = (R)(1.0 / 54.0), p3 = (R)(1.0 / 216.0), p0 = (R)(8.0 / 27.0);
    const R THREE = 3, ONE = 1, TWO = 2;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
          size_t n = I(i, j, k);
          if (std::abs((int)isf[n]) != 1) continue;
          auto F = [&](int a, int b, int c) -> R { return ld(ph, I(i + a, j + b, k + c)); };
          R gx = THREE * (p1 * (F(1, 0, 0) - F(-1, 0, 0)) +
                          p2 * ((F(1, 1, 0) - F(-1, -1, 0)) + (F(1, -1, 0) - F(-1, 1, 0)) +
                                (F(1, 0, 1) - F(-1, 0, -1)) + (F(1, 0, -1) - F(-1, 0, 1))) +
                          p3 * ((F(1, 1, 1) - F(-1, -1, -1)) + (F(1, -1, -1) - F(-1, 1, 1)) +
                                (F(1, -1, 1) - F(-1, 1, -1)) + (F(1, 1, -1) - F(-1, -1, 1))));
          R gy = THREE * (p1 * (F(0, 1, 0) - F(0, -1, 0)) +
                          p2 * ((F(1, 1, 0) - F(-1, -1, 0)) + (F(-1, 1, 0) - F(1, -1, 0)) +
                                (F(0, 1, 1) - F(0, -1, -1)) + (F(0, 1, -1) - F(0, -1, 1))) +
                          p3 * ((F(1, 1, 1) - F(-1, -1, -1)) + (F(-1, 1, -1) - F(1, -1, 1)) +
                                (F(1, 1, -1) - F(-1, -1, 1)) + (F(-1, 1, 1) - F(1, -1, -1))));
          R gz = THREE * (p1 * (F(0, 0, 1) - F(0, 0, -1)) +
                          p2 * ((F(1, 0, 1) - F(-1, 0, -1)) + (F(-1, 0, 1) - F(1, 0, -1)) +
                                (F(0, 1, 1) - F(0, -1, -1)) + (F(0, -1, 1) - F(0, 1, -1))) +
                          p3 * ((F(1, 1, 1) - F(-1, -1, -1)) + (F(-1, -1, 1) - F(1, 1, -1)) +
                                (F(1, -1, 1) - F(-1, 1, -1)) + (F(-1, 1, 1) - F(1, -1, -1))));
          R mg = std::sqrt(gx * gx + gy * gy + gz * gz);
          const R eps = (R)1.0e-9;
          R c = F(0, 0, 0);
          aux[0][n] = st(gx / (mg + eps));
          aux[1][n] = st(gy / (mg + eps));
          aux[2][n] = st(gz / (mg + eps));
          aux[3][n] = st(mg);
          aux[4][n] = st(c * (ONE - c) * (gx / (mg + eps)));
          aux[5][n] = st(c * (ONE - c) * (gy / (mg + eps)));
          aux[6][n] = st(c * (ONE - c) * (gz / (mg + eps)));
          R lap = (TWO * THREE) *
                  (c * (p0 - ONE) +
                   (p1 * (F(1, 0, 0) + F(-1, 0, 0) + F(0, 1, 0) + F(0, -1, 0) + F(0, 0, 1) + F(0, 0, -1)) +
                    p2 * ((F(1, 1, 0) + F(-1, -1, 0)) + (F(1, -1, 0) + F(-1, 1, 0)) +
                          (F(1, 0, 1) + F(-1, 0, -1)) + (F(1, 0, -1) + F(-1, 0, 1)) +
                          (F(0, 1, 1) + F(0, -1, -1)) + (F(0, 1, -1) + F(0, -1, 1))) +
                    p3 * ((F(1, 1, 1) + F(-1, -1, -1)) + (F(1, -1, -1) + F(-1, 1, 1)) +
                          (F(1, -1, 1) + F(-1, 1, -1)) + (F(1, 1, -1) + F(-1, -1, 1)))));
          loc[0][n] = st(lap);
        }
    for (int a = 0; a < 7; ++a) halo(aux[a]);  // exchange_auxfields_intpbc
  }

  // ---- compute_div_theta_n (lb_cuda_auxfields_module.f90:1128-1160) -------
  void div_theta_n() {
    const R p1 = (R)(2.0 / 27.0), p2 = (R)(1.0 / 54.0), p3 = (R)(1.0 / 216.0);
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
          size_t n = I(i, j, k);
          if (std::abs((int)isf[n]) != 1) continue;
          auto X = [&](int a, int b, int c) -> R { return ld(aux[4], I(i + a, j + b, k + c)); };
          auto Y = [&](int a, int b, int c) -> R { return ld(aux[5], I(i + a, j + b, k + c)); };
          auto Z = [&](int a, int b, int c) -> R { return ld(aux[6], I(i + a, j + b, k + c)); };
          R d = ((p1 * (X(1, 0, 0) - X(-1, 0, 0)) +
                  p2 * ((X(1, 1, 0) - X(-1, -1, 0)) + (X(1, -1, 0) - X(-1, 1, 0)) +
                        (X(1, 0, 1) - X(-1, 0, -1)) + (X(1, 0, -1) - X(-1, 0, 1))) +
                  p3 * ((X(1, 1, 1) - X(-1, -1, -1)) + (X(1, -1, -1) - X(-1, 1, 1)) +
                        (X(1, -1, 1) - X(-1, 1, -1)) + (X(1, 1, -1) - X(-1, -1, 1)))) +
                 (p1 * (Y(0, 1, 0) - Y(0, -1, 0)) +
                  p2 * ((Y(1, 1, 0) - Y(-1, -1, 0)) + (Y(-1, 1, 0) - Y(1, -1, 0)) +
                        (Y(0, 1, 1) - Y(0, -1, -1)) + (Y(0, 1, -1) - Y(0, -1, 1))) +
                  p3 * ((Y(1, 1, 1) - Y(-1, -1, -1)) + (Y(-1, 1, -1) - Y(1, -1, 1)) +
                        (Y(1, 1, -1) - Y(-1, -1, 1)) + (Y(-1, 1, 1) - Y(1, -1, -1)))) +
                 (p1 * (Z(0, 0, 1) - Z(0, 0, -1)) +
                  p2 * ((Z(1, 0, 1) - Z(-1, 0, -1)) + (Z(-1, 0, 1) - Z(1, 0, -1)) +
                        (Z(0, 1, 1) - Z(0, -1, -1)) + (Z(0, -1, 1) - Z(0, 1, -1))) +
                  p3 * ((Z(1, 1, 1) - Z(-1, -1, -1)) + (Z(-1, -1, 1) - Z(1, 1, -1)) +
                        (Z(1, -1, 1) - Z(-1, 1, -1)) + (Z(-1, 1, 1) - Z(1, -1, -1))))) *
                (R)3;
          loc[1][n] = st(d);
        }
  }

  // ---- PHI_int_boundary (lb_cuda_boundary_module.f90:274-454) -------------
  // Past a non-periodic face isfluid reads 1, so where two walls meet the
  // donor can be a node of the upper halo (the lower one is skipped, boundary:
  // 395).  The reference writes phi and the normals there nowhere: its
  // initial conditions fill the interior only (init_conditions_module.f90:
  // 60-90) and the halo exchanges copy periodic images only; here those halo
  // nodes keep the zero they are allocated with, so such a ghost becomes 0.
  void phi_int_boundary(std::vector<T>& ph) {
    const R eps_norm = (R)1.0e-14, huge_cot = (R)1.0e14;
    const R inv_sqrt2 = (R)0.70710678118654752440, inv_sqrt3 = (R)0.57735026918962576450;
    const R half_val = (R)0.5, third_val = (R)(1.0 / 3.0);
    const R pi = (R)3.1415926535897932384626433832795028841971;
    const int nl = L.nl;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
          size_t n = I(i, j, k);
          if (isf[n] != 0) continue;
          R wettab = (R)p.wettab_r;
          R ghost = ld(ph, n);
          R wnx = 0, wny = 0, wnz = 0;
          for (int l = 1; l < nl; ++l) {
            int d2 = EX[l] * EX[l] + EY[l] * EY[l] + EZ[l] * EZ[l];
            if (isf[I(i + EX[l], j + EY[l], k + EZ[l])] == 0) continue;
            R s = d2 == 1 ? (R)1 : (d2 == 2 ? half_val : third_val);
            wnx = wnx + s * (-(R)EX[l]);
            wny = wny + s * (-(R)EY[l]);
            wnz = wnz + s * (-(R)EZ[l]);
          }
          R wn = std::sqrt(wnx * wnx + wny * wny + wnz * wnz);
          if (wn < eps_norm) continue;
          wnx = wnx / wn; wny = wny / wn; wnz = wnz / wn;
          R best = (R)-1.0e30;
          int dx = 0, dy = 0, dz = 0;
          R bsx = 0, bsy = 0, bsz = 0;
          for (int l = 1; l < nl; ++l) {
            int d2 = EX[l] * EX[l] + EY[l] * EY[l] + EZ[l] * EZ[l];
            int xi = i + EX[l], yj = j + EY[l], zk = k + EZ[l];
            if (std::abs((int)isf[I(xi, yj, zk)]) != 1) continue;
            R s = d2 == 1 ? (R)1 : (d2 == 2 ? inv_sqrt2 : inv_sqrt3);
            R sx = -(R)EX[l], sy = -(R)EY[l], sz = -(R)EZ[l];
            R score = (sx * wnx + sy * wny + sz * wnz) * s;
            if (score > best) {
              best = score; dx = xi; dy = yj; dz = zk; bsx = sx; bsy = sy; bsz = sz;
            }
          }
          if (dx == 0 || dy == 0 || dz == 0) continue;
          size_t dn = I(dx, dy, dz);
          R pf = ld(ph, dn);
          R mg = ld(aux[3], dn);
          R gx = ld(aux[0], dn) * mg, gy = ld(aux[1], dn) * mg, gz = ld(aux[2], dn) * mg;
          R gdw = gx * wnx + gy * wny + gz * wnz;
          R tx = gx - gdw * wnx, ty = gy - gdw * wny, tz = gz - gdw * wnz;
          R tm = std::sqrt(tx * tx + ty * ty + tz * tz);
          R th = ((R)180 - wettab) * pi / (R)180;
          R sn = std::sin(th), cs = std::cos(th);
          R cot = std::abs(sn) < eps_norm ? std::copysign(huge_cot, cs) : cs / sn;
          R dn_ = -tm * cot;
          R bx = tx + dn_ * wnx, by = ty + dn_ * wny, bz = tz + dn_ * wnz;
          ghost = pf + (bx * bsx + by * bsy + bz * bsz) * ((R)4.0 * ghost * ((R)1.0 - ghost));
          ghost = std::max((R)0, std::min((R)1, ghost));
          ph[n] = st(ghost);
        }
  }

  // ---- moments (lb_cuda_moments_module.f90:104-339) ------------------------
  void moments(std::vector<T>* hold, std::vector<T>* hs, const std::vector<T>& ph) {
    const R cssq = (R)(1.0 / 3.0), invcssq = (R)3, HALF = (R)0.5, ONE = 1;
    const R rr = (R)p.rho_r, rb = (R)p.rho_b, beta = (R)p.beta, kapp = (R)p.kapp;
    const R v1 = (R)p.visc1, v2 = (R)p.visc2, td = (R)p.tau_diff, sc = (R)p.sharp_c;
    const R fx = (R)p.fx, fy = (R)p.fy, fz = (R)p.fz;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
          size_t n = I(i, j, k);
          if (std::abs((int)isf[n]) != 1) continue;
          R press = ld(hs[0], n);
          R phl = 0, lap = 0, mask = 0, rhophi, Fx, Fy, Fz;
          if (two) {
            phl = ld(ph, n);
            lap = ld(loc[0], n);
            mask = (R)4.0 * phl * (ONE - phl);
          }
          if (dens) {
            rhophi = rr * phl + (ONE - phl) * rb;
            if (p.buoyancy_forcing) {
              Fx = (rhophi - (rr + rb) * HALF) * fx;
              Fy = (rhophi - (rr + rb) * HALF) * fy;
              Fz = (rhophi - (rr + rb) * HALF) * fz;
            } else {
              Fx = rhophi * fx; Fy = rhophi * fy; Fz = rhophi * fz;
            }
          } else {
            rhophi = (R)1.0; Fx = fx; Fy = fy; Fz = fz;
          }
          R gfx = 0, gfy = 0, gfz = 0;
          if (two) {
            R mg = ld(aux[3], n);
            gfx = ld(aux[0], n) * mg; gfy = ld(aux[1], n) * mg; gfz = ld(aux[2], n) * mg;
            R st_ = ((R)4.0 * beta * phl * (phl - (R)1.0) * (phl - (R)0.5) - kapp * lap);
            Fx = Fx + st_ * gfx;
            Fy = Fy + st_ * gfy;
            Fz = Fz + st_ * gfz;
            if (rep) {
              // moments:148-160: clamp rho_phi * flux to +-1e-3, add it times rho_phi
              R fr[3];
              for (int c = 0; c < 3; ++c) {
                R mt = ld(loc[5 + c], n) * rhophi;
                if (std::abs(mt) > (R)1.0e-3) mt = (R)1.0e-3 * (mt < 0 ? (R)-1 : (R)1);
                fr[c] = mt * rhophi;
              }
              Fx = Fx + fr[0]; Fy = Fy + fr[1]; Fz = Fz + fr[2];
            }
          }
          R grx = 0, gry = 0, grz = 0;
          if (dens) {
            grx = (rr - rb) * gfx; gry = (rr - rb) * gfy; grz = (rr - rb) * gfz;
            Fx = Fx - mask * press * cssq * grx;
            Fy = Fy - mask * press * cssq * gry;
            Fz = Fz - mask * press * cssq * grz;
          }
          R tFx = Fx, tFy = Fy, tFz = Fz;
          R u = ld(hold[1], n), v = ld(hold[2], n), w = ld(hold[3], n);
          R visc = 0, tau = 0;
          if (dens) {
            R pxx = ld(hs[4], n), pyy = ld(hs[5], n), pzz = ld(hs[6], n);
            R pxy = ld(hs[7], n), pxz = ld(hs[8], n), pyz = ld(hs[9], n);
            pxx = pxx - u * u; pyy = pyy - v * v; pzz = pzz - w * w;
            pxy = pxy - u * v; pxz = pxz - u * w; pyz = pyz - v * w;
            visc = (rr * v1 * phl + ((R)1.0 - phl) * v2 * rb) / rhophi;
            tau = (visc * invcssq + HALF);
            Fx = Fx - mask * (visc / (tau * cssq)) * (pxx * grx + pxy * gry + pxz * grz);
            Fy = Fy - mask * (visc / (tau * cssq)) * (pyy * gry + pxy * grx + pyz * grz);
            Fz = Fz - mask * (visc / (tau * cssq)) * (pzz * grz + pxz * grx + pyz * gry);
          }
          u = ld(hs[1], n); v = ld(hs[2], n); w = ld(hs[3], n);
          if (incomp) {
            R t = -sc * ld(loc[1], n);
            u = u + (R)0.5 * Fx / (rhophi);
            v = v + (R)0.5 * Fy / (rhophi);
            w = w + (R)0.5 * Fz / (rhophi);
            R den = ((R)1.0 - (R)0.5 * (rr - rb) * (td * lap + t) / rhophi);
            u = u / den;
            v = v / den;
            w = w / den;
            R a = (rr - rb) * (td * lap + t);
            tFx = tFx - a * u; tFy = tFy - a * v; tFz = tFz - a * w;
            Fx = Fx - a * u; Fy = Fy - a * v; Fz = Fz - a * w;
          } else {
            u = u + (R)0.5 * Fx / rhophi;
            v = v + (R)0.5 * Fy / rhophi;
            w = w + (R)0.5 * Fz / rhophi;
          }
          hs[1][n] = st(u); hs[2][n] = st(v); hs[3][n] = st(w);
          R pxx = ld(hs[4], n), pyy = ld(hs[5], n), pzz = ld(hs[6], n);
          R pxy = ld(hs[7], n), pxz = ld(hs[8], n), pyz = ld(hs[9], n);
          pxx = pxx + Fx * u / rhophi;
          pyy = pyy + Fy * v / rhophi;
          pzz = pzz + Fz * w / rhophi;
          pxy = pxy + HALF * (Fy * u + Fx * v) / rhophi;
          pxz = pxz + HALF * (Fz * u + Fx * w) / rhophi;
          pyz = pyz + HALF * (Fz * v + Fy * w) / rhophi;
          hs[4][n] = st(pxx); hs[5][n] = st(pyy); hs[6][n] = st(pzz);
          hs[7][n] = st(pxy); hs[8][n] = st(pxz); hs[9][n] = st(pyz);
          if (dens) {
            pxx = pxx - u * u; pyy = pyy - v * v; pzz = pzz - w * w;
            pxy = pxy - u * v; pxz = pxz - u * w; pyz = pyz - v * w;
            tFx = tFx - mask * (visc / (tau * cssq)) * (pxx * grx + pxy * gry + pxz * grz);
            tFy = tFy - mask * (visc / (tau * cssq)) * (pyy * gry + pxy * grx + pyz * grz);
            tFz = tFz - mask * (visc / (tau * cssq)) * (pzz * grz + pxz * grx + pyz * gry);
          }
          frc[0][n] = st(tFx); frc[1][n] = st(tFy); frc[2][n] = st(tFz);
        }
    for (int f = 0; f < 10; ++f) halo(hs[f]);  // exchange_hfields_intpbc
    for (int f = 0; f < 3; ++f) halo(frc[f]);  // exchange_forces_intpbc
  }

  // ---- post-collision populations of one node (fused:530-713) ----------------
  void collide_node(size_t n, const std::vector<T>* hin, const std::vector<T>& ph, R* f) const {
    const R cssq = (R)(1.0 / 3.0), invcssq = (R)3, HALF = (R)0.5, ONE = 1, TWO = 2, THREE = 3;
    const R FOUR = 4, SIX = 6, SEVEN = 7, TEN = 10;
    R phl = 0, inv;
    if (two) phl = ld(ph, n);
    if (dens) inv = (R)1.0 / ((R)p.rho_r * phl + (ONE - phl) * (R)p.rho_b);
    else inv = (R)1.0;
    R fx = ld(frc[0], n) * inv, fy = ld(frc[1], n) * inv, fz = ld(frc[2], n) * inv;
    R press = ld(hin[0], n), u = ld(hin[1], n), v = ld(hin[2], n), w = ld(hin[3], n);
    R pxx = ld(hin[4], n), pyy = ld(hin[5], n), pzz = ld(hin[6], n);
    R pxy = ld(hin[7], n), pxz = ld(hin[8], n), pyz = ld(hin[9], n);
    if (obst && isf[n] == 0) {
      fx = fy = fz = 0; press = u = v = w = 0; pxx = pyy = pzz = pxy = pxz = pyz = 0;
    }
    pxx = pxx - u * u; pyy = pyy - v * v; pzz = pzz - w * w;
    pxy = pxy - u * v; pxz = pxz - u * w; pyz = pyz - v * w;
    R om;
    const R visc = two ? ((R)p.rho_r * (R)p.visc1 * phl + (ONE - phl) * (R)p.visc2 * (R)p.rho_b) * inv : (R)p.visc1;
    if (smag) {
      om = smag_omega(visc, pxx, pyy, pzz, pxy, pxz, pyz, inv);
    } else if (two) {
      om = (visc * invcssq + (R)0.5);
      om = (R)1.0 / om;
    } else {
      om = (R)p.omega;
    }
    R uu = HALF * (u * u + v * v + w * w) * invcssq;
    // l = 0
    R feq = (R)L.w[0] * (press - uu);
    R fneq = (R)L.w[0] * (HALF * invcssq) * (-pxx - pyy - pzz);
    if (ho) {
      feq = feq + Minv[17][0] * u * u * v * v + Minv[18][0] * u * u * w * w + Minv[19][0] * v * v * w * w +
            Minv[26][0] * u * u * v * v * w * w;
      fneq = fneq + Minv[17][0] * (pyy * u * u + TWO * v * (THREE * pxy * u + pxx * v)) +
             Minv[18][0] * (pzz * u * u + TWO * w * (THREE * pxz * u + pxx * w)) +
             Minv[19][0] * (pzz * v * v + TWO * w * (THREE * pyz * v + pyy * w)) +
             Minv[26][0] * (pzz * u * u * v * v +
                            w * (SIX * u * v * (pyz * u + pxz * v) + THREE * pyy * u * u * w +
                                 TWO * v * (SEVEN * pxy * u + TWO * pxx * v) * w));
    }
    R Fd = (R)L.w[0] * (-u * fx - v * fy - w * fz) * invcssq;
    f[0] = feq + (ONE - om) * fneq + HALF * Fd;
    for (int l = 1; l < L.nl; ++l) {
      const R ex = (R)EX[l], ey = (R)EY[l], ez = (R)EZ[l];
      R ud = (u * ex + v * ey + w * ez) * invcssq;
      feq = (R)L.w[l] * (press + (ud + (R)0.5 * ud * ud - uu));
      fneq = (R)L.w[l] * (HALF / (cssq * cssq)) *
             ((ex * ex - cssq) * pxx + (ey * ey - cssq) * pyy + (ez * ez - cssq) * pzz + TWO * (ex * ey) * pxy +
              TWO * (ex * ez) * pxz + TWO * (ey * ez) * pyz);
      if (ho) {
        feq = feq + Minv[10][l] * u * u * v + Minv[11][l] * u * u * w + Minv[12][l] * u * v * v +
              Minv[13][l] * u * w * w + Minv[14][l] * v * w * w + Minv[15][l] * v * v * w +
              Minv[16][l] * u * v * w + Minv[17][l] * u * u * v * v + Minv[18][l] * u * u * w * w +
              Minv[19][l] * v * v * w * w + Minv[20][l] * u * v * w * w + Minv[21][l] * u * v * v * w +
              Minv[22][l] * u * u * v * w + Minv[23][l] * u * u * v * w * w + Minv[24][l] * u * u * v * v * w +
              Minv[25][l] * u * v * v * w * w + Minv[26][l] * u * u * v * v * w * w;
        fneq = fneq + Minv[10][l] * (TWO * pxy * u + pxx * v) + Minv[11][l] * (TWO * pxz * u + pxx * w) +
               Minv[12][l] * (pyy * u + TWO * pxy * v) + Minv[13][l] * (pzz * u + TWO * pxz * w) +
               Minv[14][l] * (pzz * v + TWO * pyz * w) + Minv[15][l] * (TWO * pyz * v + pyy * w) +
               Minv[16][l] * (pyz * u + pxz * v + pxy * w) +
               Minv[17][l] * (pyy * u * u + TWO * v * (THREE * pxy * u + pxx * v)) +
               Minv[18][l] * (pzz * u * u + TWO * w * (THREE * pxz * u + pxx * w)) +
               Minv[19][l] * (pzz * v * v + TWO * w * (THREE * pyz * v + pyy * w)) +
               Minv[20][l] * (pzz * u * v + w * (THREE * pyz * u + THREE * pxz * v + TWO * pxy * w)) +
               Minv[21][l] * (TWO * pyz * u * v + pxz * v * v + TWO * pyy * u * w + FOUR * pxy * v * w) +
               Minv[22][l] * (pyz * u * u + TWO * pxz * u * v + FOUR * pxy * u * w + TWO * pxx * v * w) +
               Minv[23][l] * (pzz * u * u * v + THREE * w * (pyz * u * u + TWO * pxz * u * v + TWO * pxy * u * w + pxx * v * w)) +
               Minv[24][l] * (TWO * u * v * (pyz * u + pxz * v) + TWO * pyy * u * u * w + v * (TEN * pxy * u + THREE * pxx * v) * w) +
               Minv[25][l] * (pzz * u * v * v + THREE * w * (TWO * pyz * u * v + pxz * v * v + pyy * u * w + TWO * pxy * v * w)) +
               Minv[26][l] * (pzz * u * u * v * v +
                              w * (SIX * u * v * (pyz * u + pxz * v) + THREE * pyy * u * u * w +
                                   TWO * v * (SEVEN * pxy * u + TWO * pxx * v) * w));
      }
      Fd = (R)L.w[l] * (((ex - u) + ud * ex) * fx + ((ey - v) + ud * ey) * fy + ((ez - w) + ud * ez) * fz) * invcssq;
      f[l] = feq + (ONE - om) * fneq + HALF * Fd;
    }
  }

  // ---- fused (fused:494-748): collide every node of the halo grid, push ------
  void fused(const std::vector<T>* hin, std::vector<T>* hout, const std::vector<T>& ph) {
    const int nl = L.nl;
    pop.assign(N * 27, (R)0);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < hz; ++k)
      for (int j = 0; j < hy; ++j)
        for (int i = 0; i < hx; ++i) {
          size_t n = I(i, j, k);
          collide_node(n, hin, ph, &pop[n * 27]);
        }
    const R cssq = (R)(1.0 / 3.0);
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
          size_t n = I(i, j, k);
          R f0 = pop[n * 27 + 0];
          R op = 0, ou = 0, ov = 0, ow = 0, oxx = 0, oyy = 0, ozz = 0, oxy = 0, oxz = 0, oyz = 0;
          op = op + f0;
          oxx = oxx + f0 * (-cssq);
          oyy = oyy + f0 * (-cssq);
          ozz = ozz + f0 * (-cssq);
          for (int l = 1; l < nl; ++l) {
            const R ex = (R)EX[l], ey = (R)EY[l], ez = (R)EZ[l];
            const size_t up = I(i - EX[l], j - EY[l], k - EZ[l]);
            R f = pop[up * 27 + l];
            // BOUNCE_BACK (fused:364-382): the solid upstream node hands the
            // node's own opposite population back (fluid nodes; solid-node
            // outputs are unused and order-dependent in the reference)
            if (bb && isf[n] != 0 && isf[up] == 0) f = pop[n * 27 + OPP[l]];
            op = op + f;
            ou = ou + f * ex; ov = ov + f * ey; ow = ow + f * ez;
            oxx = oxx + f * (ex * ex - cssq);
            oyy = oyy + f * (ey * ey - cssq);
            ozz = ozz + f * (ez * ez - cssq);
            oxy = oxy + f * ex * ey; oxz = oxz + f * ex * ez; oyz = oyz + f * ey * ez;
          }
          hout[0][n] = st(op); hout[1][n] = st(ou); hout[2][n] = st(ov); hout[3][n] = st(ow);
          hout[4][n] = st(oxx); hout[5][n] = st(oyy); hout[6][n] = st(ozz);
          hout[7][n] = st(oxy); hout[8][n] = st(oxz); hout[9][n] = st(oyz);
        }
  }

  // ---- update_phi (lb_cuda_update_phi_module.f90:77-101) -----------------------
  void update_phi(const std::vector<T>* hs, const std::vector<T>& pin, std::vector<T>& pout) {
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
          size_t n = I(i, j, k);
          if (std::abs((int)isf[n]) != 1) continue;
          R t = -(R)p.sharp_c * ld(loc[1], n);
          R mg = ld(aux[3], n);
          R gx = ld(aux[0], n) * mg, gy = ld(aux[1], n) * mg, gz = ld(aux[2], n) * mg;
          R ph = ld(pin, n), lap = ld(loc[0], n);
          R u = ld(hs[1], n), v = ld(hs[2], n), w = ld(hs[3], n);
          R o = ph - u * (R)0.5 * (gx) - v * (R)0.5 * (gy) - w * (R)0.5 * (gz) + (R)p.tau_diff * lap + t;
          pout[n] = st(o);
        }
    halo(pout);  // exchange_phifields_intpbc
  }

  // ---- LB_int_boundary (lb_cuda_boundary_module.f90:85-212) ---------------------
  void lb_int_boundary(const std::vector<T>* hin, std::vector<T>* hout, const std::vector<T>& ph) {
    const R cssq = (R)(1.0 / 3.0), invcssq = (R)3, HALF = (R)0.5, ONE = 1, TWO = 2;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
          size_t n = I(i, j, k);
          if (isf[n] != -1) continue;
          R phl = 0, rhophi = 1;
          if (two) phl = ld(ph, n);
          if (dens) rhophi = (R)p.rho_r * phl + (ONE - phl) * (R)p.rho_b;
          R press = ld(hin[0], n), u = ld(hin[1], n), v = ld(hin[2], n), w = ld(hin[3], n);
          R pxx = ld(hin[4], n), pyy = ld(hin[5], n), pzz = ld(hin[6], n);
          R pxy = ld(hin[7], n), pxz = ld(hin[8], n), pyz = ld(hin[9], n);
          R op = ld(hout[0], n), ou = ld(hout[1], n), ov = ld(hout[2], n), ow = ld(hout[3], n);
          R oxx = ld(hout[4], n), oyy = ld(hout[5], n), ozz = ld(hout[6], n);
          R oxy = ld(hout[7], n), oxz = ld(hout[8], n), oyz = ld(hout[9], n);
          pxx = pxx - u * u; pyy = pyy - v * v; pzz = pzz - w * w;
          pxy = pxy - u * v; pxz = pxz - u * w; pyz = pyz - v * w;
          R om;
          const R visc = two ? ((R)p.rho_r * (R)p.visc1 * phl + (ONE - phl) * (R)p.visc2 * (R)p.rho_b) / rhophi
                             : (R)p.visc1;
          if (smag) {
            // boundary:150-155 divides by rho_phi where fused multiplies by its inverse
            const R QQ = pxx * pxx + pyy * pyy + pzz * pzz + TWO * (pxy * pxy + pxz * pxz + pyz * pyz);
            const R C = (R)(double)(0.053f * 18.0f);
            const R t3 = (R)3.0 * visc;
            om = (R)0.5 + ((R)1.0 / (R)6.0) * ((R)3.0 * visc + std::sqrt(t3 * t3 + C * std::sqrt((R)2.0 * QQ) / rhophi));
            om = (R)1.0 / om;
          } else if (two) {
            om = (visc * invcssq + (R)0.5);
            om = (R)1.0 / om;
          } else {
            om = (R)p.omega;
          }
          for (int l = 1; l < L.nl; ++l) {
            int lo = OPP[l];
            if (isf[I(i + EX[lo], j + EY[lo], k + EZ[lo])] != 0) continue;
            const R ex = (R)EX[l], ey = (R)EY[l], ez = (R)EZ[l];
            R feq = (R)L.w[l] * (press);
            R fneq = (R)L.w[l] * (HALF / (cssq * cssq)) *
                     ((ex * ex - cssq) * pxx + (ey * ey - cssq) * pyy + (ez * ez - cssq) * pzz +
                      TWO * (ex * ey) * pxy + TWO * (ex * ez) * pxz + TWO * (ey * ez) * pyz);
            R f = feq + (ONE - om) * fneq;
            op = op + f;
            ou = ou + f * ex; ov = ov + f * ey; ow = ow + f * ez;
            oxx = oxx + f * (ex * ex - cssq);
            oyy = oyy + f * (ey * ey - cssq);
            ozz = ozz + f * (ez * ez - cssq);
            oxy = oxy + f * ex * ey; oxz = oxz + f * ex * ez; oyz = oyz + f * ey * ez;
          }
          hout[0][n] = st(op); hout[1][n] = st(ou); hout[2][n] = st(ov); hout[3][n] = st(ow);
          hout[4][n] = st(oxx); hout[5][n] = st(oyy); hout[6][n] = st(ozz);
          hout[7][n] = st(oxy); hout[8][n] = st(oxz); hout[9][n] = st(oyz);
        }
  }

  // ---- REPULSIVE_FLUX (lb_cuda_repulsive_module.f90) --------------------------------
  // Neighbours up to 3 nodes away wrap periodically (the reference's 3-deep
  // halo copies, mpi_module.f90:163-165); across a non-periodic face they are
  // skipped.
  bool wrap3(int& i, int& j, int& k) const {
    auto w = [](int& a, int n, int per) {
      if (a >= 1 && a <= n) return true;
      if (!per) return false;
      a = ((a - 1) % n + n) % n + 1;
      return true;
    };
    return w(i, nx, p.pbc[0]) && w(j, ny, p.pbc[1]) && w(k, nz, p.pbc[2]);
  }

  // thinfilm_scan_mark_kernel (repulsive:19-150)
  void thinfilm_scan(const std::vector<T>& ph) {
    const int win = 3;  // LBFAST.f90:221
    const R q_th = (R)p.q_th, cosOpp = (R)p.cos_opp_t;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
          size_t n = I(i, j, k);
          if (std::abs((int)isf[n]) != 1) continue;
          rep_mask[n] = 0;
          loc[2][n] = T(0); loc[3][n] = T(0); loc[4][n] = T(0);
          R q = std::min(std::max(ld(ph, n), (R)0), (R)1);
          q = q * ((R)1 - q);
          if (q < q_th) continue;
          const R nix = ld(aux[0], n), niy = ld(aux[1], n), niz = ld(aux[2], n);
          R best_r2 = std::numeric_limits<R>::max(), best_face = (R)-1;
          int bi = 0, bj = 0, bk = 0;
          bool found = false;
          for (int di = -win; di <= win; ++di)
            for (int dj = -win; dj <= win; ++dj)
              for (int dk = -win; dk <= win; ++dk) {
                if (di == 0 && dj == 0 && dk == 0) continue;
                int a = i + di, b = j + dj, c = k + dk;
                if (!wrap3(a, b, c)) continue;
                const size_t m = I(a, b, c);
                if (std::abs((int)isf[m]) != 1) continue;
                const R r2 = (R)di * (R)di + (R)dj * (R)dj + (R)dk * (R)dk;
                if (r2 < (R)1.0e-12) continue;
                R qn = std::min(std::max(ld(ph, m), (R)0), (R)1);
                qn = qn * ((R)1 - qn);
                if (qn < q_th || std::abs(qn - q) > (R)0.1 * std::max(q, (R)1.0e-12)) continue;
                const R dotn = nix * ld(aux[0], m) + niy * ld(aux[1], m) + niz * ld(aux[2], m);
                if (dotn > cosOpp) continue;
                const R face = (R)0.5 * ((R)1 - dotn);
                if (r2 < best_r2 - (R)1.0e-14) {
                  best_r2 = r2; best_face = face; bi = i + di; bj = j + dj; bk = k + dk; found = true;
                } else if (std::abs(r2 - best_r2) <= (R)1.0e-14) {
                  if (face > best_face) { best_face = face; bi = i + di; bj = j + dj; bk = k + dk; found = true; }
                }
              }
          if (found) {
            loc[2][n] = (T)(R)bi; loc[3][n] = (T)(R)bj; loc[4][n] = (T)(R)bk;
            rep_mask[n] = 1;
          }
        }
  }

  // repulsive_flux_normal_kernel (repulsive:152-289)
  void repulsive_flux(const std::vector<T>& ph) {
    const R eps = (R)1.0e-9, width = (R)p.width, A_rep = (R)p.a_rep;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
          size_t n = I(i, j, k);
          if (std::abs((int)isf[n]) != 1) continue;
          loc[5][n] = T(0); loc[6][n] = T(0); loc[7][n] = T(0);
          if (rep_mask[n] != 1) continue;
          const R phl = ld(ph, n);
          const R q1 = phl * ((R)1 - phl);
          if (q1 <= eps) continue;
          const int pi = (int)ld(loc[2], n), pj = (int)ld(loc[3], n), pk = (int)ld(loc[4], n);
          R dx = (R)(pi - i), dy = (R)(pj - j), dz = (R)(pk - k);
          const R r = std::sqrt(dx * dx + dy * dy + dz * dz);
          if (r <= eps) continue;
          const R rinv = (R)1.0 / r;
          dx = dx * rinv; dy = dy * rinv; dz = dz * rinv;
          int a = pi, b = pj, c = pk;
          if (!wrap3(a, b, c)) continue;
          const size_t m = I(a, b, c);
          const R nx1 = ld(aux[0], n), ny1 = ld(aux[1], n), nz1 = ld(aux[2], n);
          const R nx2 = ld(aux[0], m), ny2 = ld(aux[1], m), nz2 = ld(aux[2], m);
          const R face = std::max((R)0, -(nx1 * nx2 + ny1 * ny2 + nz1 * nz2));
          if (face <= eps) continue;
          R nsx = nx1 - nx2, nsy = ny1 - ny2, nsz = nz1 - nz2;
          const R nsmag = std::sqrt(nsx * nsx + nsy * nsy + nsz * nsz);
          if (nsmag <= eps) {
            nsx = dx; nsy = dy; nsz = dz;
          } else {
            nsx = nsx / nsmag; nsy = nsy / nsmag; nsz = nsz / nsmag;
          }
          if (dx * nsx + dy * nsy + dz * nsz < (R)0) { nsx = -nsx; nsy = -nsy; nsz = -nsz; }
          const R phl2 = ld(ph, m);
          const R q2 = phl2 * ((R)1 - phl2);
          const R qpair = (R)0.5 * (q1 + q2);
          const R qcl = std::min(std::max(qpair, eps), (R)0.25 - eps);
          const R arg = (R)1.0 / ((R)2.0 * std::sqrt(qcl));
          if (arg <= (R)1.0) continue;
          const R ach = std::log(arg + std::sqrt(arg * arg - (R)1.0));
          const R Wfilm = width * ach;
          if (Wfilm <= (R)0) continue;
          const R wdth = (R)1.0 / ((R)1.0 + std::pow(Wfilm, (R)4.0));
          const R fx = A_rep * wdth * qcl * face * nsx;
          const R fy = A_rep * wdth * qcl * face * nsy;
          const R fz = A_rep * wdth * qcl * face * nsz;
          const R cap = (R)1.5 * (std::abs(fx) + std::abs(fy) + std::abs(fz));
          const R scales = std::min((R)1.0, phl / std::max(cap, (R)1.0e-9));
          loc[5][n] = st(fx * scales); loc[6][n] = st(fy * scales); loc[7][n] = st(fz * scales);
        }
  }

  // ---- open boundaries (boundary_cds_module.f90:57-1311, 1353-1664) ------------------
  // Face f = 2*axis + side (x(1) x(2) y(1) y(2) z(1) z(2)) of each non-periodic
  // axis; the loops over the face run over the interior with the face's edges
  // skipped (gj > 1 .and. gj < ly, ...), in the reference's face order.
  template <class Fn>
  void face_loop(int f, int layer_lo, int layer_hi, Fn fn) const {
    const int ax = f >> 1, side = f & 1;
    const int n3[3] = {nx, ny, nz};
    if (p.pbc[ax]) return;
    const int b = ax == 0 ? 1 : 0, c = ax == 2 ? 1 : 2;  // the two other axes
    const int lay = side ? layer_hi : layer_lo;
    for (int q = 2; q <= n3[c] - 1; ++q)
      for (int r = 2; r <= n3[b] - 1; ++r) {
        int ijk[3];
        ijk[ax] = lay; ijk[b] = r; ijk[c] = q;
        fn(ijk[0], ijk[1], ijk[2]);
      }
  }
  // the HIGHORDER terms of the regularised population of link l (the E_k and
  // N_k of fused:668-702 with Minv_mrt(10..26, l)), as the open-face loops write them
  R ho_feq(int l, R u, R v, R w) const {
    return Minv[10][l] * u * u * v + Minv[11][l] * u * u * w + Minv[12][l] * u * v * v +
           Minv[13][l] * u * w * w + Minv[14][l] * v * w * w + Minv[15][l] * v * v * w +
           Minv[16][l] * u * v * w + Minv[17][l] * u * u * v * v + Minv[18][l] * u * u * w * w +
           Minv[19][l] * v * v * w * w + Minv[20][l] * u * v * w * w + Minv[21][l] * u * v * v * w +
           Minv[22][l] * u * u * v * w + Minv[23][l] * u * u * v * w * w + Minv[24][l] * u * u * v * v * w +
           Minv[25][l] * u * v * v * w * w + Minv[26][l] * u * u * v * v * w * w;
  }
  R ho_fneq(int l, R pxx, R pyy, R pzz, R pxy, R pxz, R pyz, R u, R v, R w) const {
    const R TWO = 2, THREE = 3, FOUR = 4, SIX = 6, SEVEN = 7, TEN = 10;
    return Minv[10][l] * (TWO * pxy * u + pxx * v) + Minv[11][l] * (TWO * pxz * u + pxx * w) +
           Minv[12][l] * (pyy * u + TWO * pxy * v) + Minv[13][l] * (pzz * u + TWO * pxz * w) +
           Minv[14][l] * (pzz * v + TWO * pyz * w) + Minv[15][l] * (TWO * pyz * v + pyy * w) +
           Minv[16][l] * (pyz * u + pxz * v + pxy * w) +
           Minv[17][l] * (pyy * u * u + TWO * v * (THREE * pxy * u + pxx * v)) +
           Minv[18][l] * (pzz * u * u + TWO * w * (THREE * pxz * u + pxx * w)) +
           Minv[19][l] * (pzz * v * v + TWO * w * (THREE * pyz * v + pyy * w)) +
           Minv[20][l] * (pzz * u * v + w * (THREE * pyz * u + THREE * pxz * v + TWO * pxy * w)) +
           Minv[21][l] * (TWO * pyz * u * v + pxz * v * v + TWO * pyy * u * w + FOUR * pxy * v * w) +
           Minv[22][l] * (pyz * u * u + TWO * pxz * u * v + FOUR * pxy * u * w + TWO * pxx * v * w) +
           Minv[23][l] * (pzz * u * u * v + THREE * w * (pyz * u * u + TWO * pxz * u * v + TWO * pxy * u * w + pxx * v * w)) +
           Minv[24][l] * (TWO * u * v * (pyz * u + pxz * v) + TWO * pyy * u * u * w + v * (TEN * pxy * u + THREE * pxx * v) * w) +
           Minv[25][l] * (pzz * u * v * v + THREE * w * (TWO * pyz * u * v + pxz * v * v + pyy * u * w + TWO * pxy * v * w)) +
           Minv[26][l] * (pzz * u * u * v * v +
                          w * (SIX * u * v * (pyz * u + pxz * v) + THREE * pyy * u * u * w + TWO * v * (SEVEN * pxy * u + TWO * pxx * v) * w));
  }
  // bcs_mesoscopic_phifields (1353-1664): the wall node takes phi of its
  // neighbour inside the domain
  void open_phi(std::vector<T>& ph) {
    for (int f = 0; f < 6; ++f) {
      const int ax = f >> 1, d = (f & 1) ? -1 : 1;
      face_loop(f, 1, ax == 0 ? nx : (ax == 1 ? ny : nz), [&](int i, int j, int k) {
        ph[I(i, j, k)] = ph[I(i + (ax == 0) * d, j + (ax == 1) * d, k + (ax == 2) * d)];
      });
    }
    halo(ph);
  }
  // bcs_mesoscopic_hfields (57-1311), after LB_int_boundary: on the first
  // fluid layer, the links from solid nodes lose the closure's population
  // built from the node's P and Pi - cs^2 P - u u and gain the regularised
  // population of the face state (presstmp, utmp, vtmp, wtmp by openbc_type)
  void open_h(const std::vector<T>* hin, std::vector<T>* hout, const std::vector<T>& ph) {
    const R cssq = (R)(1.0 / 3.0), invcssq = (R)3, HALF = (R)0.5, ONE = 1, TWO = 2;
    for (int f = 0; f < 6; ++f) {
      const int ax = f >> 1;
      const int nA = ax == 0 ? nx : (ax == 1 ? ny : nz);
      face_loop(f, 2, nA - 1, [&](int i, int j, int k) {
        const size_t n = I(i, j, k);
        if (isf[n] != -1) return;
        R press = ld(hin[0], n), u = ld(hin[1], n), v = ld(hin[2], n), w = ld(hin[3], n);
        R pxx = ld(hin[4], n), pyy = ld(hin[5], n), pzz = ld(hin[6], n);
        R pxy = ld(hin[7], n), pxz = ld(hin[8], n), pyz = ld(hin[9], n);
        R op = ld(hout[0], n), ou = ld(hout[1], n), ov = ld(hout[2], n), ow = ld(hout[3], n);
        R oxx = ld(hout[4], n), oyy = ld(hout[5], n), ozz = ld(hout[6], n);
        R oxy = ld(hout[7], n), oxz = ld(hout[8], n), oyz = ld(hout[9], n);
        R presstmp = press, utmp = u, vtmp = v, wtmp = w;
        if (p.openbc_type[f] == 1) {
          presstmp = (R)p.openbc_press[f];
        } else if (p.openbc_type[f] != 2) {
          utmp = (R)p.openbc_u[f]; vtmp = (R)p.openbc_v[f]; wtmp = (R)p.openbc_w[f];
        }
        pxx = pxx - cssq * press - u * u;
        pyy = pyy - cssq * press - v * v;
        pzz = pzz - cssq * press - w * w;
        pxy = pxy - u * v;
        pxz = pxz - u * w;
        pyz = pyz - v * w;
        R om;
        if (two) {
          const R phl = ld(ph, n);
          const R rhophi = dens ? (R)p.rho_r * phl + (ONE - phl) * (R)p.rho_b : ONE;
          const R visc = ((R)p.rho_r * (R)p.visc1 * phl + ((R)1.0 - phl) * (R)p.visc2 * (R)p.rho_b) / rhophi;
          const R tau = (visc * invcssq + HALF);
          om = ONE / tau;
        } else {
          om = (R)p.omega;
        }
        const R uu = HALF * (utmp * utmp + vtmp * vtmp + wtmp * wtmp) * invcssq;
        for (int l = 1; l < L.nl; ++l) {
          const int lo = OPP[l];
          if (isf[I(i + EX[lo], j + EY[lo], k + EZ[lo])] != 0) continue;
          const R ex = (R)EX[l], ey = (R)EY[l], ez = (R)EZ[l], wl = (R)L.w[l];
          R feq = wl * (press);
          R fneq1 = wl * (HALF / (cssq * cssq)) *
                    ((ex * ex - cssq) * pxx + (ey * ey - cssq) * pyy + (ez * ez - cssq) * pzz +
                     TWO * (ex * ey) * pxy + TWO * (ex * ez) * pxz + TWO * (ey * ez) * pyz);
          R fpost = feq + (ONE - om) * fneq1;
          op = op - fpost;
          ou = ou - fpost * ex; ov = ov - fpost * ey; ow = ow - fpost * ez;
          oxx = oxx - fpost * ex * ex; oyy = oyy - fpost * ey * ey; ozz = ozz - fpost * ez * ez;
          oxy = oxy - fpost * ex * ey; oxz = oxz - fpost * ex * ez; oyz = oyz - fpost * ey * ez;
          const R udotc = (utmp * ex + vtmp * ey + wtmp * ez) * invcssq;
          feq = wl * (presstmp + udotc + HALF * udotc * udotc - uu);
          fneq1 = wl * (HALF / (cssq * cssq)) *
                  ((ex * ex - cssq) * pxx + (ey * ey - cssq) * pyy + (ez * ez - cssq) * pzz +
                   TWO * (ex * ey) * pxy + TWO * (ex * ez) * pxz + TWO * (ey * ez) * pyz);
          if (ho) {
            feq = feq + ho_feq(l, utmp, vtmp, wtmp);
            fneq1 = fneq1 + ho_fneq(l, pxx, pyy, pzz, pxy, pxz, pyz, utmp, vtmp, wtmp);
          }
          fpost = feq + (ONE - om) * fneq1;
          op = op + fpost;
          ou = ou + fpost * ex; ov = ov + fpost * ey; ow = ow + fpost * ez;
          oxx = oxx + fpost * (ex * ex - cssq); oyy = oyy + fpost * (ey * ey - cssq);
          ozz = ozz + fpost * (ez * ez - cssq);
          oxy = oxy + fpost * ex * ey; oxz = oxz + fpost * ex * ez; oyz = oyz + fpost * ey * ez;
        }
        hout[0][n] = st(op); hout[1][n] = st(ou); hout[2][n] = st(ov); hout[3][n] = st(ow);
        hout[4][n] = st(oxx); hout[5][n] = st(oyy); hout[6][n] = st(ozz);
        hout[7][n] = st(oxy); hout[8][n] = st(oxz); hout[9][n] = st(oyz);
      });
    }
  }
  // BCPHIFLUX with open boundaries (1667-1862): the phase carried through the
  // first fluid layer of each open face, + u_a phi at the low face and
  // - u_a phi at the high one, u the velocity along the face's axis in S;
  // every node of the layer with isfluid -1, edges included
  double open_flux(const std::vector<T>* hs, const std::vector<T>& ph) const {
    double chg = 0;
    const int n3[3] = {nx, ny, nz};
    for (int f = 0; f < 6; ++f) {
      const int ax = f >> 1, side = f & 1;
      if (p.pbc[ax]) continue;
      const int lay = side ? n3[ax] - 1 : 2;
      for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
          for (int i = 1; i <= nx; ++i) {
            const int ijk[3] = {i, j, k};
            if (ijk[ax] != lay) continue;
            const size_t n = I(i, j, k);
            if (isf[n] != -1) continue;
            const double ua = (double)hs[1 + ax][n], phl = (double)ph[n];
            chg = side ? chg - ua * phl : chg + ua * phl;
          }
    }
    return chg;
  }

  // ---- BCPHIFLUX phase-mass correction ------------------------------------------
  // phi_sum_count_kernel / apply_lagrangian_phi_kernel (lb_cuda_boundary_module.f90:
  // 461-566) as boundary_cds_module.f90:1864-1884 drives them, without open
  // boundaries (global_phi_change = 0): the phi sum over |isfluid| == 1 nodes
  // (double), the count of those with 0.5 < phi < 0.9, corr = (sum_ini - sum) /
  // real(count) with the count as a default (single precision) real, added to
  // the counted nodes at storage precision.  The reference refreshes no phi
  // halo afterwards, so its fused stage would collide periodic halo images
  // with the uncorrected value; here the halo is refreshed, as an in-kernel
  // periodic wrap reads the corrected value.
  void phi_sums(const std::vector<T>& ph, double& sum, long& cnt) const {
    sum = 0;
    cnt = 0;
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
          const size_t n = I(i, j, k);
          if (std::abs((int)isf[n]) != 1) continue;
          const double v = (double)ph[n];
          sum += v;
          if (v > 0.5 && v < 0.9) ++cnt;
        }
  }
  void phi_lagrange(std::vector<T>& ph) {
    double sum;
    long cnt;
    phi_sums(ph, sum, cnt);
    // dphi = (global_phi_sum_ini + global_phi_change) - global_phi_sum, the
    // change being this step's flux through the open faces (zero without them)
    const double chg = obc ? open_flux(h[cur], ph) : 0.0;
    const double corr = cnt > 0 ? ((phi_sum_ini + chg) - sum) / (double)(float)cnt : 0.0;
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
          const size_t n = I(i, j, k);
          if (std::abs((int)isf[n]) != 1) continue;
          const double v = (double)ph[n];
          if (v > 0.5 && v < 0.9) ph[n] = (T)(v + corr);
        }
    halo(ph);
  }

  // ---- driver-level entry points ---------------------------------------------------
  int begin_step() override { ++stepno; return 0; }
  int end_step() override { cur = 1 - cur; return 0; }

  int stage(int id) override {
    std::vector<T>* S = h[cur];
    std::vector<T>* O = h[1 - cur];
    std::vector<T>& pS = phi[cur];
    std::vector<T>& pO = phi[1 - cur];
    switch (id) {
      case LBF_STAGE_NORM: if (!two) return 0; compute_norm(pS); return 0;
      case LBF_STAGE_PHI_INT_BOUNDARY: if (!two || !obst) return 0; phi_int_boundary(pS); return 0;
      case LBF_STAGE_DIV_THETA_N: if (!two) return 0; div_theta_n(); return 0;
      case LBF_STAGE_MOMENTS: moments(O, S, pS); return 0;
      case LBF_STAGE_FUSED: fused(S, O, pS); return 0;
      case LBF_STAGE_UPDATE_PHI: if (!two) return 0; update_phi(S, pS, pO); return 0;
      case LBF_STAGE_LB_INT_BOUNDARY: if (!obst || bb) return 0; lb_int_boundary(S, O, pS); return 0;
      case LBF_STAGE_THINFILM_SCAN: if (!rep) return 0; thinfilm_scan(pS); return 0;
      case LBF_STAGE_REPULSIVE_FLUX: if (!rep) return 0; repulsive_flux(pS); return 0;
      case LBF_STAGE_PHI_LAGRANGE: if (!lagr) return 0; phi_lagrange(pS); return 0;
      case LBF_STAGE_OPEN_BC_PHI: if (!two || !obc) return 0; open_phi(pS); return 0;
      case LBF_STAGE_OPEN_BC_H: if (!obc) return 0; open_h(S, O, pS); return 0;
      default: err = "unknown stage"; return LBF_EINVAL;
    }
  }

  // integrator_module.f90:266-545 (one FLIP half; FLOP is the same with S/O swapped)
  int step(int nsteps) override {
    for (int s = 0; s < nsteps; ++s) {
      begin_step();
      stage(LBF_STAGE_NORM);
      stage(LBF_STAGE_PHI_INT_BOUNDARY);
      stage(LBF_STAGE_OPEN_BC_PHI);
      stage(LBF_STAGE_PHI_LAGRANGE);
      stage(LBF_STAGE_DIV_THETA_N);
      stage(LBF_STAGE_THINFILM_SCAN);
      stage(LBF_STAGE_REPULSIVE_FLUX);
      stage(LBF_STAGE_MOMENTS);
      stage(LBF_STAGE_FUSED);
      stage(LBF_STAGE_UPDATE_PHI);
      stage(LBF_STAGE_LB_INT_BOUNDARY);
      stage(LBF_STAGE_OPEN_BC_H);
      end_step();
    }
    return 0;
  }

  // init_conditions_module.f90:584-587 (O := S) and integrator_module.f90:179-233
  int init_state() override {
    for (int f = 0; f < 10; ++f) h[1 - cur][f] = h[cur][f];
    phi[1 - cur] = phi[cur];
    for (auto& a : aux) std::fill(a.begin(), a.end(), T(0));
    if (lagr) {  // global_phi_sum_ini (init_conditions_module.f90:469-490)
      long cnt;
      phi_sums(phi[cur], phi_sum_ini, cnt);
    }
    if (two && obst) phi_int_boundary(phi[cur]);  // bcs_phi before the loop
    if (two && obc) open_phi(phi[cur]);
    if (lagr) phi_lagrange(phi[cur]);
    if (two) halo(phi[cur]);
    for (int f = 0; f < 10; ++f) halo(h[cur][f]);
    // the copy O := S happened before the halos were filled; fill O's too so
    // that solid/halo reads of O are the same finite values
    for (int f = 0; f < 10; ++f) halo(h[1 - cur][f]);
    if (two) halo(phi[1 - cur]);
    stepno = 0;
    return 0;
  }

  std::vector<T>* field(int id) {
    if (id >= 1 && id <= 10) return &h[cur][id - 1];
    if (id == LBF_FIELD_PHI) return &phi[cur];
    if (id >= 12 && id <= 14) return &h[1 - cur][id - 11];  // u_old lives in O(2:4)
    if (id >= 15 && id <= 17) return &frc[id - 15];
    if (id >= 18 && id <= 24) return &aux[id - 18];
    if (id == 25 || id == 26) return &loc[id - 25];
    if (id >= 41 && id <= 46) return &loc[id - 39];
    if (id >= 27 && id <= 36) return &h[1 - cur][id - 27];
    if (id == LBF_FIELD_PHI_O) return &phi[1 - cur];
    return nullptr;
  }

  int set_field(int id, const double* src) override {
    auto* a = field(id);
    if (!a) { err = "bad field id"; return LBF_EINVAL; }
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i)
          (*a)[I(i, j, k)] = (T)(R)src[(size_t)(i - 1) + (size_t)nx * ((j - 1) + (size_t)ny * (k - 1))];
    halo(*a);
    return 0;
  }
  int get_field(int id, double* dst) override {
    if (id == LBF_FIELD_REP_MASK) {
      for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
          for (int i = 1; i <= nx; ++i)
            dst[(size_t)(i - 1) + (size_t)nx * ((j - 1) + (size_t)ny * (k - 1))] = (double)rep_mask[I(i, j, k)];
      return 0;
    }
    if (id == LBF_FIELD_RHOPHI) {
      for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
          for (int i = 1; i <= nx; ++i) {
            R ph = two ? ld(phi[cur], I(i, j, k)) : (R)0;
            R v = dens ? (R)p.rho_r * ph + ((R)1 - ph) * (R)p.rho_b : (two ? ph : ld(h[cur][0], I(i, j, k)));
            dst[(size_t)(i - 1) + (size_t)nx * ((j - 1) + (size_t)ny * (k - 1))] = (double)v;
          }
      return 0;
    }
    auto* a = field(id);
    if (!a) { err = "bad field id"; return LBF_EINVAL; }
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i)
          dst[(size_t)(i - 1) + (size_t)nx * ((j - 1) + (size_t)ny * (k - 1))] = (double)(*a)[I(i, j, k)];
    return 0;
  }
  int set_isfluid(const int8_t* src) override {
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i)
          isf[I(i, j, k)] = src[(size_t)(i - 1) + (size_t)nx * ((j - 1) + (size_t)ny * (k - 1))];
    // halo: periodic image, else 1 (LBFAST.f90:258 sets the whole array to 1)
    auto wrap = [](int i, int n) { return i < 1 ? i + n : (i > n ? i - n : i); };
    for (int k = 0; k < hz; ++k)
      for (int j = 0; j < hy; ++j)
        for (int i = 0; i < hx; ++i) {
          bool ox = i == 0 || i == hx - 1, oy = j == 0 || j == hy - 1, oz = k == 0 || k == hz - 1;
          if (!(ox || oy || oz)) continue;
          if ((ox && !p.pbc[0]) || (oy && !p.pbc[1]) || (oz && !p.pbc[2])) { isf[I(i, j, k)] = 1; continue; }
          isf[I(i, j, k)] = isf[I(wrap(i, nx), wrap(j, ny), wrap(k, nz))];
        }
    return 0;
  }
  // post-collision populations of the last FUSED stage, interior nodes, [l][k][j][i]
  int populations(double* dst) override {
    if (pop.empty()) { err = "no fused stage run yet"; return LBF_ESTATE; }
    const size_t V = (size_t)nx * ny * nz;
    for (int l = 0; l < 27; ++l)
      for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
          for (int i = 1; i <= nx; ++i)
            dst[l * V + (size_t)(i - 1) + (size_t)nx * ((j - 1) + (size_t)ny * (k - 1))] = (double)pop[I(i, j, k) * 27 + l];
    return 0;
  }
};

}  // namespace orc

// ---------------------------------------------------------------- C ABI for ctypes
extern "C" {

void* orc_create(const lbf_params* p) {
  if (!p || p->nx < 1 || p->ny < 1 || p->nz < 1) return nullptr;
  if (p->lattice != 19 && p->lattice != 27) return nullptr;
  switch (p->precision) {
    case LBF_PREC_FP64: return new orc::Sim<double, double>(*p);
    case LBF_PREC_FP32: return new orc::Sim<float, float>(*p);
    case LBF_PREC_MIXED: return new orc::Sim<double, float>(*p);
    default: return nullptr;
  }
}
void orc_destroy(void* h) { delete (orc::SimBase*)h; }
int orc_set_field(void* h, int id, const double* src) { return ((orc::SimBase*)h)->set_field(id, src); }
int orc_get_field(void* h, int id, double* dst) { return ((orc::SimBase*)h)->get_field(id, dst); }
int orc_set_isfluid(void* h, const int8_t* src) { return ((orc::SimBase*)h)->set_isfluid(src); }
int orc_init_state(void* h) { return ((orc::SimBase*)h)->init_state(); }
int orc_begin_step(void* h) { return ((orc::SimBase*)h)->begin_step(); }
int orc_stage(void* h, int id) { return ((orc::SimBase*)h)->stage(id); }
int orc_end_step(void* h) { return ((orc::SimBase*)h)->end_step(); }
int orc_step(void* h, int n) { return ((orc::SimBase*)h)->step(n); }
int orc_populations(void* h, double* dst) { return ((orc::SimBase*)h)->populations(dst); }
int orc_classify_isfluid(const lbf_params* p, const int8_t* in, int8_t* out) { return orc::classify(p, in, out); }
void orc_derive(lbf_params* p) { orc::derive(p); }
double orc_minv(int k, int l) { return orc::minv_d3q27(k, l); }
double orc_m(int k, int l) { return orc::m_d3q27(k, l); }
// Wall-clock seconds of n steps (the CPU baseline timer).
double orc_time_steps(void* h, int n) {
  auto t0 = std::chrono::steady_clock::now();
  ((orc::SimBase*)h)->step(n);
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}
}
