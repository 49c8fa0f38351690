// lbf_common.h -- device-side definitions shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

namespace lbf {

// ---------------------------------------------------------------------------
// Kernel parameter block (by value).  Doubles are converted to the compute
// type R at kernel entry.
struct DevParams {
  int nx, ny, nz;
  int pbc[3];
  int twocomp, dens, incomp, obst, buoy;
  int lattice, highorder;
  double fx, fy, fz;
  double visc1, visc2, rho_r, rho_b;
  double tau_diff, sharp_c, beta, kapp, omega;
  double wettab;
  float scale[10];     // FP16S/BF16S: stored = scale * deviation
  float invscale[10];
};

__host__ __device__ inline int wrapi(int i, int n) {
  // i in [-n, 2n)
  return i < 0 ? i + n : (i >= n ? i - n : i);
}
__host__ __device__ inline size_t nidx(int x, int y, int z, int nx, int ny) {
  return (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * (size_t)z);
}

// ---------------------------------------------------------------------------
// Storage codecs.  A codec maps the compute type R <-> the storage type T.
// Plain codecs store the value; the deviation codecs (FP16S/BF16S) store
// scale*value for P and u, and scale*(Pi_ab - u_a u_b) for the second
// moments, so that all stored quantities are small deviations from the rest
// state and land in the normal range of the 16-bit format.
enum Prec { PREC_FP64 = 0, PREC_FP32 = 1, PREC_MIXED = 2, PREC_FP16S = 3, PREC_BF16S = 4 };

template <int P> struct Codec;
template <> struct Codec<PREC_FP64> {
  typedef double T; typedef double R;
  static constexpr bool dev = false;
  __device__ static R ld(const T* a, size_t i, float) { return a[i]; }
  __device__ static void stv(T* a, size_t i, R v, float) { a[i] = v; }
  __device__ static R round(R v) { return v; }
};
template <> struct Codec<PREC_FP32> {
  typedef float T; typedef float R;
  static constexpr bool dev = false;
  __device__ static R ld(const T* a, size_t i, float) { return a[i]; }
  __device__ static void stv(T* a, size_t i, R v, float) { a[i] = v; }
  __device__ static R round(R v) { return v; }
};
template <> struct Codec<PREC_MIXED> {
  typedef float T; typedef double R;
  static constexpr bool dev = false;
  __device__ static R ld(const T* a, size_t i, float) { return (R)a[i]; }
  __device__ static void stv(T* a, size_t i, R v, float) { a[i] = (T)v; }
  // the reference stores intermediates (S after moments, forces) in strdb
  __device__ static R round(R v) { return (R)(T)v; }
};
template <> struct Codec<PREC_FP16S> {
  typedef __half T; typedef float R;
  static constexpr bool dev = true;
  __device__ static R ld(const T* a, size_t i, float inv) { return __half2float(a[i]) * inv; }
  __device__ static void stv(T* a, size_t i, R v, float s) { a[i] = __float2half_rn(v * s); }
  __device__ static R round(R v) { return v; }
};
template <> struct Codec<PREC_BF16S> {
  typedef __hip_bfloat16 T; typedef float R;
  static constexpr bool dev = true;
  __device__ static R ld(const T* a, size_t i, float inv) { return __bfloat162float(a[i]) * inv; }
  __device__ static void stv(T* a, size_t i, R v, float s) { a[i] = __float2bfloat16(v * s); }
  __device__ static R round(R v) { return v; }
};

// phase field / aux / forces / u_old storage type per precision mode
template <int P> struct Aux { typedef typename Codec<P>::T T; };
template <> struct Aux<PREC_FP16S> { typedef float T; };
template <> struct Aux<PREC_BF16S> { typedef float T; };
template <> struct Aux<PREC_MIXED> { typedef float T; };

// The ten moments of one node in compute precision.
template <class R> struct Mom {
  R P, u, v, w, xx, yy, zz, xy, xz, yz;
};

// Load/store the ten moment fields (SoA pointers m[0..9]).  For the
// deviation codecs the second moments are stored as Pi - u u, with u taken
// after its own rounding so that a load reproduces Pi consistently.
template <int P>
__device__ inline Mom<typename Codec<P>::R> load_mom(typename Codec<P>::T* const* m, size_t i, const DevParams& dp) {
  typedef Codec<P> C; typedef typename C::R R;
  Mom<R> o;
  o.P = C::ld(m[0], i, dp.invscale[0]);
  o.u = C::ld(m[1], i, dp.invscale[1]);
  o.v = C::ld(m[2], i, dp.invscale[2]);
  o.w = C::ld(m[3], i, dp.invscale[3]);
  o.xx = C::ld(m[4], i, dp.invscale[4]);
  o.yy = C::ld(m[5], i, dp.invscale[5]);
  o.zz = C::ld(m[6], i, dp.invscale[6]);
  o.xy = C::ld(m[7], i, dp.invscale[7]);
  o.xz = C::ld(m[8], i, dp.invscale[8]);
  o.yz = C::ld(m[9], i, dp.invscale[9]);
  if (C::dev) {
    o.xx += o.u * o.u; o.yy += o.v * o.v; o.zz += o.w * o.w;
    o.xy += o.u * o.v; o.xz += o.u * o.w; o.yz += o.v * o.w;
  }
  return o;
}

template <int P>
__device__ inline void store_mom(typename Codec<P>::T* const* m, size_t i, const Mom<typename Codec<P>::R>& o,
                                 const DevParams& dp) {
  typedef Codec<P> C; typedef typename C::R R;
  C::stv(m[0], i, o.P, dp.scale[0]);
  if (C::dev) {
    // round u to storage first, then store Pi - u_r u_r
    typename C::T tu, tv, tw;
    tu = (typename C::T)0; tv = tu; tw = tu;
    C::stv(&tu, 0, o.u, dp.scale[1]);
    C::stv(&tv, 0, o.v, dp.scale[2]);
    C::stv(&tw, 0, o.w, dp.scale[3]);
    R ur = C::ld(&tu, 0, dp.invscale[1]), vr = C::ld(&tv, 0, dp.invscale[2]), wr = C::ld(&tw, 0, dp.invscale[3]);
    m[1][i] = tu; m[2][i] = tv; m[3][i] = tw;
    C::stv(m[4], i, o.xx - ur * ur, dp.scale[4]);
    C::stv(m[5], i, o.yy - vr * vr, dp.scale[5]);
    C::stv(m[6], i, o.zz - wr * wr, dp.scale[6]);
    C::stv(m[7], i, o.xy - ur * vr, dp.scale[7]);
    C::stv(m[8], i, o.xz - ur * wr, dp.scale[8]);
    C::stv(m[9], i, o.yz - vr * wr, dp.scale[9]);
  } else {
    C::stv(m[1], i, o.u, 0.f); C::stv(m[2], i, o.v, 0.f); C::stv(m[3], i, o.w, 0.f);
    C::stv(m[4], i, o.xx, 0.f); C::stv(m[5], i, o.yy, 0.f); C::stv(m[6], i, o.zz, 0.f);
    C::stv(m[7], i, o.xy, 0.f); C::stv(m[8], i, o.xz, 0.f); C::stv(m[9], i, o.yz, 0.f);
  }
}

// ---------------------------------------------------------------------------
// Lattice.  Velocity ordering of the reference (vars_module.f90:319-338; the
// first 19 are the D3Q19 set, 666-684).  Kernels address populations by
// their velocity (p,q,r) in {-1,0,1}^3 through cube index c = (p+1) +
// 3(q+1) + 9(r+1).
__host__ __device__ constexpr int cidx(int p, int q, int r) { return (p + 1) + 3 * (q + 1) + 9 * (r + 1); }

template <int LAT>
__host__ __device__ constexpr bool present(int p, int q, int r) {
  // D3Q27: all 27; D3Q19: drop the 8 corners (|p|+|q|+|r| == 3)
  return LAT == 27 ? true : ((p != 0 ? 1 : 0) + (q != 0 ? 1 : 0) + (r != 0 ? 1 : 0)) < 3;
}
template <int LAT>
__host__ __device__ constexpr double weight(int p, int q, int r) {
  return LAT == 27
             ? ((p == 0 ? 2.0 / 3.0 : 1.0 / 6.0) * (q == 0 ? 2.0 / 3.0 : 1.0 / 6.0) * (r == 0 ? 2.0 / 3.0 : 1.0 / 6.0))
             : (((p != 0) + (q != 0) + (r != 0)) == 0 ? 1.0 / 3.0
                : (((p != 0) + (q != 0) + (r != 0)) == 1 ? 1.0 / 18.0 : 1.0 / 36.0));
}

// One-axis Hermite reconstruction of the tensor-product D3Q27 basis:
// f(c) = sum_a g_a(c) A_a with g_0 = w1d(c), g_1 = 3 w1d c,
// g_2 = 4.5 w1d (c^2 - 1/3), w1d = (1/6, 2/3, 1/6).  (Minv_mrt of
// vars_module.f90:359-521 is the product of three of these.)
template <class R>
__device__ __forceinline__ void recon3(R A0, R A1, R A2, R& fm, R& f0, R& fp) {
  const R t = A0 * (R)(1.0 / 6.0) + A2 * (R)0.5;
  fp = t + A1 * (R)0.5;
  fm = t - A1 * (R)0.5;
  f0 = A0 * (R)(2.0 / 3.0) - A2;
}

}  // namespace lbf
