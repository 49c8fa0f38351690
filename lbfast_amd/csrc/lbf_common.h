// lbf_common.h -- device-side definitions shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

namespace lbf {

// Physical scalars in the compute type (converted once on the host).
template <class R>
struct Phys {
  R fx, fy, fz;
  R visc1, visc2, rho_r, rho_b;
  R tau_diff, sharp_c, beta, kapp, omega;
  R wettab;
  R width, a_rep, q_th, cos_opp;  // REPULSIVE_FLUX
};

// Kernel parameter block (by value).
struct DevParams {
  int nx, ny, nz;
  int pbc[3];
  int twocomp, dens, incomp, obst, buoy;
  int lattice, highorder;
  int bb, smag, rep;     // BOUNCE_BACK (with obstacles), SMAGORINSKI, REPULSIVE_FLUX
  size_t nn;             // nx*ny*nz nodes
  int sx;                // row length of one field in a field set (nx rounded up to even)
  Phys<double> pd;
  Phys<float> pf;
  float scale[10];       // FP16S/BF16S: stored = scale * deviation
  float invscale[10];
  template <class R> __device__ __forceinline__ const Phys<R>& ph() const;
};
template <> __device__ __forceinline__ const Phys<double>& DevParams::ph<double>() const { return pd; }
template <> __device__ __forceinline__ const Phys<float>& DevParams::ph<float>() const { return pf; }

// Single fields (phi, lap, div theta, masks) are x-fastest arrays of nx*ny*nz.
__host__ __device__ inline size_t nidx(int x, int y, int z, int nx, int ny) {
  return (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * (size_t)z);
}

// Field sets (the ten moments, u_old, forces, aux, the prep output) are
// row-interleaved: the F field rows of one (y, z) row follow each other, each
// sx values long, so element (x, y, z) of field f sits at
//   ((z*ny + y)*F + f)*sx + x.
// A wave's 64-lane segment of every field of a set then lies within one
// F*sx run of memory instead of F arrays nx*ny*nz apart, which keeps the
// streaming kernels' DRAM pages and translations local.  sbase() is field 0;
// field f adds f*sx.
__host__ __device__ inline size_t sbase(int F, int x, int y, int z, int sx, int ny) {
  return ((size_t)z * ny + y) * (size_t)F * sx + x;
}
__host__ __device__ inline int set_sx(int nx) { return (nx + 1) & ~1; }
__host__ __device__ inline size_t set_elems(int F, int nx, int ny, int nz) {
  return (size_t)F * set_sx(nx) * ny * nz;
}

// Load through a wave-uniform base pointer and a 32-bit byte offset (selects
// the scalar-base + vector-offset global addressing mode).
template <class T>
__device__ __forceinline__ T ldb(const T* base, uint32_t boff) {
  return *(const T*)((const char*)base + boff);
}
template <class T>
__device__ __forceinline__ void stb(T* base, uint32_t boff, T v) {
  *(T*)((char*)base + boff) = v;
}

// Buffer-resource access: a wave-uniform base (scalar registers) plus a
// 32-bit per-lane byte offset, range-checked against num_records (an offset
// past the plane reads 0 / drops the store instead of faulting).  Resource
// word 3 is the CDNA (gfx9) raw-buffer format.
typedef __amdgpu_buffer_rsrc_t Rsrc;
__device__ __forceinline__ Rsrc mkrsrc(const void* p, uint32_t nbytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)nbytes, 0x00020000);
}
template <class T> __device__ __forceinline__ T bld(Rsrc r, uint32_t boff);
template <> __device__ __forceinline__ float bld<float>(Rsrc r, uint32_t boff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, boff, 0, 0));
}
template <> __device__ __forceinline__ double bld<double>(Rsrc r, uint32_t boff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, boff, 0, 0));
}
template <> __device__ __forceinline__ __half bld<__half>(Rsrc r, uint32_t boff) {
  return __builtin_bit_cast(__half, (unsigned short)__builtin_amdgcn_raw_buffer_load_b16(r, boff, 0, 0));
}
template <> __device__ __forceinline__ __hip_bfloat16 bld<__hip_bfloat16>(Rsrc r, uint32_t boff) {
  return __builtin_bit_cast(__hip_bfloat16, (unsigned short)__builtin_amdgcn_raw_buffer_load_b16(r, boff, 0, 0));
}
template <> __device__ __forceinline__ int8_t bld<int8_t>(Rsrc r, uint32_t boff) {
  return (int8_t)__builtin_amdgcn_raw_buffer_load_b8(r, boff, 0, 0);
}
__device__ __forceinline__ void bst(Rsrc r, uint32_t boff, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, boff, 0, 0);
}
__device__ __forceinline__ void bst(Rsrc r, uint32_t boff, double v) {
  typedef decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)) V2;
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, boff, 0, 0);
}
__device__ __forceinline__ void bst(Rsrc r, uint32_t boff, __half v) {
  __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), r, boff, 0, 0);
}
__device__ __forceinline__ void bst(Rsrc r, uint32_t boff, __hip_bfloat16 v) {
  __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), r, boff, 0, 0);
}
// field f of a field set at one z plane
template <class T>
__device__ __forceinline__ T fld(const T* plane0, size_t fs, int f, uint32_t pbytes, uint32_t boff) {
  return bld<T>(mkrsrc(plane0 + (size_t)f * fs, pbytes), boff);
}
template <class T>
__device__ __forceinline__ void fst(T* plane0, size_t fs, int f, uint32_t pbytes, uint32_t boff, T v) {
  bst(mkrsrc(plane0 + (size_t)f * fs, pbytes), boff, v);
}

// LDS DMA.  buffer_load_dword with the lds modifier writes each lane's dword
// to LDS at M0 + 4*lane instead of to a VGPR: a wave can stream the next
// plane's inputs into its own LDS slots while it computes, at no register
// cost.  Issued from inline asm so that the compiler's wait-count tracking
// (which would otherwise make every later s_barrier wait for the transfer)
// does not see it; the consumer waits with dma_wait_all() before reading.
// M0 (= lds + OFF, OFF a compile-time byte offset) is saved and restored
// around the load, and s_nop 4 covers the M0-write and SGPR-write hazards
// ahead of the memory instruction.
typedef __attribute__((address_space(3))) float lds_float;
__device__ __forceinline__ uint32_t lds_addr(float* p) {
  return __builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)(lds_float*)p);
}
template <uint32_t OFF = 0>
__device__ __forceinline__ void dma_dword(Rsrc r, uint32_t voff, uint32_t lds) {
  uint32_t save;
  asm volatile(
      "s_mov_b32 %0, m0\n\t"
      "s_add_u32 m0, %3, %4\n\t"
      "s_nop 4\n\t"
      "buffer_load_dword %1, %2, 0 offen lds\n\t"
      "s_mov_b32 m0, %0"
      : "=&s"(save)
      : "v"(voff), "s"(r), "s"(lds), "n"(OFF)
      : "memory", "scc");
}
// Wait for every outstanding vector-memory operation of this wave (the
// counter also covers stores, so issue a plane's stores before its DMA).
__device__ __forceinline__ void dma_wait_all() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
__device__ __forceinline__ void lds_wait_all() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// Separable D3Q27 stencils.  The isotropic weights of the gradient and
// Laplacian stencils (lb_cuda_auxfields_module.f90:72-130) are
// w(a,b,c) = u_a u_b u_c / 216 with u = (1, 4, 1) over offsets (-1, 0, 1), so
//   DX(F) = sum_c u_c sum_b u_b (F(1,b,c) - F(-1,b,c)) / 216   (DY, DZ alike)
//   LAPL  = 6 (sum_abc w(a,b,c) F(a,b,c) - F(0,0,0)).
// One z plane contributes three 9-point partial sums per node; the z
// direction combines the partials of three consecutive planes m, 0, p.
// The 2-D smoothing s is taken of the differences to the plane's own centre
// value c (which is kept beside it): summing the 36-weight smoothing of O(1)
// values and cancelling afterwards, as s_p - s_m and s_m + 4 s_0 + s_p - 216 c
// do, loses about six bits in FP32 at a smooth interface; differences first
// keep the gradient as accurate as the reference's paired differences and the
// Laplacian more accurate than its weighted sum.  With S = s + 36 c:
//   grad = 3/216 (dx_m + 4 dx_0 + dx_p,  dy_m + 4 dy_0 + dy_p,  s_p - s_m) + (0, 0, (c_p - c_m)/2)
//   lap  = 6/216 (s_m + 4 s_0 + s_p) + (c_m - c_0) + (c_p - c_0)
// the partial sums of one plane at one node
template <class R> struct Part3 {
  R dx, dy;  // x difference smoothed in y, y difference smoothed in x
  R s, c;    // 2-D smoothing of (value - c), and c = the value at the node's column
};

// partials of the 3 x 3 neighbourhood v[b][a] (b: y offset + 1, a: x offset + 1)
template <class R>
__device__ __forceinline__ Part3<R> part3(const R (&v)[3][3]) {
  const R FOUR = 4;
  Part3<R> q;
  q.dx = (v[0][2] - v[0][0]) + FOUR * (v[1][2] - v[1][0]) + (v[2][2] - v[2][0]);
  q.dy = (v[2][0] - v[0][0]) + FOUR * (v[2][1] - v[0][1]) + (v[2][2] - v[0][2]);
  const R c = v[1][1];
  const R r0 = (v[0][0] - c) + FOUR * (v[0][1] - c) + (v[0][2] - c);
  const R r1 = (v[1][0] - c) + (v[1][2] - c);
  const R r2 = (v[2][0] - c) + FOUR * (v[2][1] - c) + (v[2][2] - c);
  q.s = r0 + FOUR * r1 + r2;
  q.c = c;
  return q;
}

// gradient and Laplacian at a node from the partials of planes z-1, z, z+1
// (m, o, p); o.c is the node's own value
template <class R>
__device__ __forceinline__ void grad_lap(const Part3<R>& m, const Part3<R>& o, const Part3<R>& p, R& gx, R& gy,
                                         R& gz, R& lap) {
  const R FOUR = 4, HALF = 0.5, G3 = (R)(3.0 / 216.0), L6 = (R)(6.0 / 216.0);
  gx = G3 * (m.dx + FOUR * o.dx + p.dx);
  gy = G3 * (m.dy + FOUR * o.dy + p.dy);
  gz = G3 * (p.s - m.s) + HALF * (p.c - m.c);
  lap = L6 * (m.s + FOUR * o.s + p.s) + ((m.c - o.c) + (p.c - o.c));
}

// XCD-aware tile order.  Workgroups are dispatched round-robin over the 8
// XCDs (workgroup b runs on XCD b % 8), each with its own L2.  Give XCD k a
// contiguous range of tiles in (x fastest, y, z) order, so the tiles resident
// together on one XCD are x- and y-neighbours that share halo cache lines.
constexpr int kNumXCD = 8;
__device__ __forceinline__ int xcd_tile(int b, int ntiles) {
  const int per = ntiles / kNumXCD, rem = ntiles % kNumXCD;
  const int k = b % kNumXCD, slot = b / kNumXCD;
  return k * per + (k < rem ? k : rem) + slot;
}

// Reciprocal: 1-ulp hardware estimate in FP32 (well inside the FP32
// tolerance), correctly rounded division in FP64.
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp(double x) { return 1.0 / x; }

// ---------------------------------------------------------------------------
// Storage codecs: compute type R <-> storage type T.  Plain codecs store the
// value; the deviation codecs (FP16S/BF16S) store scale*value for P and u and
// scale*(Pi_ab - u_a u_b) for the second moments, so every stored quantity is
// a small deviation from the rest state inside the 16-bit normal range.
enum Prec { PREC_FP64 = 0, PREC_FP32 = 1, PREC_MIXED = 2, PREC_FP16S = 3, PREC_BF16S = 4 };

// Stochastic rounding of the 16-bit codecs.  Round-to-nearest drops every
// per-step change smaller than half a 16-bit ulp (11 significant bits), so
// a slowly driven flow stalls and drifts; rounding up with probability equal
// to the dropped fraction makes each store unbiased in expectation.  The
// random word of a field is the bits of another field of the same node
// (its FP32 value, whose low bits are chaotic from step to step) XOR a
// per-node, per-plane, per-step seed from a 32-bit integer hash (the murmur3
// finaliser).  No state is kept between launches, and replays of a captured
// graph draw new bits because the values change from step to step.
__device__ __forceinline__ uint32_t mix32(uint32_t h) {
  h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
  return h;
}

template <int P> struct Codec;
template <> struct Codec<PREC_FP64> {
  typedef double T; typedef double R;
  static constexpr bool dev = false;
  __device__ static R dec(T v, float) { return v; }
  __device__ static T enc(R v, float) { return v; }
  __device__ static R round(R v) { return v; }
};
template <> struct Codec<PREC_FP32> {
  typedef float T; typedef float R;
  static constexpr bool dev = false;
  __device__ static R dec(T v, float) { return v; }
  __device__ static T enc(R v, float) { return v; }
  __device__ static R round(R v) { return v; }
};
template <> struct Codec<PREC_MIXED> {
  typedef float T; typedef double R;
  static constexpr bool dev = false;
  __device__ static R dec(T v, float) { return (R)v; }
  __device__ static T enc(R v, float) { return (T)v; }
  // the reference stores intermediates (S after moments, forces) in strdb
  __device__ static R round(R v) { return (R)(T)v; }
};
template <> struct Codec<PREC_FP16S> {
  typedef __half T; typedef float R;
  static constexpr bool dev = true;
  __device__ static R dec(T v, float inv) { return __half2float(v) * inv; }
  // a second moment from its stored deviation: v * inv + uu, one mixed-
  // precision FMA (the half operand is widened inside the instruction)
  __device__ static R decpi(T v, float inv, R uu) { return __builtin_fmaf(__half2float(v), inv, uu); }
  __device__ static T enc(R v, float s) { return __float2half_rn(v * s); }
  // stochastic rounding of the scaled value x: r holds 13 random bits (a
  // uniform offset in [0, 1) ulp of the half significand, in the normal range
  // where every stored scaled deviation sits), added to the magnitude, then a
  // conversion toward zero, which drops the bits below the half significand:
  // the value rounds away from zero with probability equal to the dropped
  // fraction.  The sign bit is untouched (the add cannot carry into it for
  // finite x), and x = 0 stays 0.  Below the half normal range the dither
  // covers less than an ulp, so tiny values lean toward zero.  (gfx950's
  // v_cvt_sr_f16_f32 rounds stochastically in one instruction but made the
  // fused kernel 3 % slower.)
  static constexpr uint32_t RBITS = 0x1FFFu;
  __device__ static T srx(float x, uint32_t r) {
    const auto h2 = __builtin_amdgcn_cvt_pkrtz(__uint_as_float(__float_as_uint(x) + r), 0.0f);
    return __builtin_bit_cast(T, h2[0]);
  }
  __device__ static R round(R v) { return v; }
};
template <> struct Codec<PREC_BF16S> {
  typedef __hip_bfloat16 T; typedef float R;
  static constexpr bool dev = true;
  __device__ static R dec(T v, float inv) { return __bfloat162float(v) * inv; }
  __device__ static R decpi(T v, float inv, R uu) { return __builtin_fmaf(__bfloat162float(v), inv, uu); }
  __device__ static T enc(R v, float s) { return __float2bfloat16(v * s); }
  // stochastic rounding of the scaled value x: r holds 16 random bits, added
  // below the bfloat16 significand, then truncation
  static constexpr uint32_t RBITS = 0xFFFFu;
  __device__ static T srx(float x, uint32_t r) {
    return __builtin_bit_cast(T, (unsigned short)((__float_as_uint(x) + r) >> 16));
  }
  __device__ static R round(R v) { return v; }
};

// phase field / aux / forces storage type per precision mode
template <int P> struct Aux { typedef typename Codec<P>::T T; };
template <> struct Aux<PREC_FP16S> { typedef float T; };
template <> struct Aux<PREC_BF16S> { typedef float T; };
template <> struct Aux<PREC_MIXED> { typedef float T; };

// The 16-bit codecs' default scales: powers of two that put typical
// deviations (P ~ 1e-4..1e-2, u ~ 1e-3..1e-1, Pi - u u ~ 1e-7..1e-3) in the
// normal range of binary16.  A kernel instantiated for PREC | PREC_DEFSCALE
// has them compiled in as instruction literals; with the scales as kernel
// arguments the twenty values occupy scalar registers, which the fused
// kernel then spills to vector-register lanes and reloads every plane.
constexpr int PREC_DEFSCALE = 8;
__host__ __device__ constexpr int prec_base(int P) { return P & 7; }
__host__ __device__ constexpr float def_scale(int i) { return i == 0 ? 16384.0f : (i < 4 ? 1024.0f : 1048576.0f); }
template <> struct Codec<PREC_FP16S | PREC_DEFSCALE> : Codec<PREC_FP16S> {};
template <> struct Codec<PREC_BF16S | PREC_DEFSCALE> : Codec<PREC_BF16S> {};
template <> struct Aux<PREC_FP16S | PREC_DEFSCALE> { typedef float T; };
template <> struct Aux<PREC_BF16S | PREC_DEFSCALE> { typedef float T; };
// scale and inverse scale of field i (i a constant after unrolling)
template <int P>
__device__ __forceinline__ float scl(const DevParams& dp, int i) {
  if constexpr ((P & PREC_DEFSCALE) != 0) return def_scale(i);
  else return dp.scale[i];
}
template <int P>
__device__ __forceinline__ float iscl(const DevParams& dp, int i) {
  if constexpr ((P & PREC_DEFSCALE) != 0) return 1.0f / def_scale(i);
  else return dp.invscale[i];
}

// The ten moments of one node in compute precision.
template <class R> struct Mom {
  R P, u, v, w, xx, yy, zz, xy, xz, yz;
};

// Ten-field moment set: field f at base + f*fs (fs = sx for the row-interleaved
// sets).  `base` already includes any wave-uniform plane offset; `boff` is
// the per-lane byte offset.
template <int P>
__device__ __forceinline__ Mom<typename Codec<P>::R> load_mom(const typename Codec<P>::T* base, size_t fs,
                                                              uint32_t boff, const DevParams& dp) {
  typedef Codec<P> C; typedef typename C::R R;
  Mom<R> o;
  o.P = C::dec(ldb(base + 0 * fs, boff), dp.invscale[0]);
  o.u = C::dec(ldb(base + 1 * fs, boff), dp.invscale[1]);
  o.v = C::dec(ldb(base + 2 * fs, boff), dp.invscale[2]);
  o.w = C::dec(ldb(base + 3 * fs, boff), dp.invscale[3]);
  o.xx = C::dec(ldb(base + 4 * fs, boff), dp.invscale[4]);
  o.yy = C::dec(ldb(base + 5 * fs, boff), dp.invscale[5]);
  o.zz = C::dec(ldb(base + 6 * fs, boff), dp.invscale[6]);
  o.xy = C::dec(ldb(base + 7 * fs, boff), dp.invscale[7]);
  o.xz = C::dec(ldb(base + 8 * fs, boff), dp.invscale[8]);
  o.yz = C::dec(ldb(base + 9 * fs, boff), dp.invscale[9]);
  if constexpr (C::dev) {
    o.xx = C::decpi(ldb(base + 4 * fs, boff), dp.invscale[4], o.u * o.u);
    o.yy = C::decpi(ldb(base + 5 * fs, boff), dp.invscale[5], o.v * o.v);
    o.zz = C::decpi(ldb(base + 6 * fs, boff), dp.invscale[6], o.w * o.w);
    o.xy = C::decpi(ldb(base + 7 * fs, boff), dp.invscale[7], o.u * o.v);
    o.xz = C::decpi(ldb(base + 8 * fs, boff), dp.invscale[8], o.u * o.w);
    o.yz = C::decpi(ldb(base + 9 * fs, boff), dp.invscale[9], o.v * o.w);
  }
  return o;
}

template <int P>
__device__ __forceinline__ void store_mom(typename Codec<P>::T* base, size_t fs, uint32_t boff,
                                          const Mom<typename Codec<P>::R>& o, const DevParams& dp) {
  typedef Codec<P> C; typedef typename C::R R; typedef typename C::T T;
  stb(base + 0 * fs, boff, C::enc(o.P, dp.scale[0]));
  if (C::dev) {
    // round u first, then store Pi - u_r u_r so a load reproduces Pi
    const T tu = C::enc(o.u, dp.scale[1]), tv = C::enc(o.v, dp.scale[2]), tw = C::enc(o.w, dp.scale[3]);
    const R ur = C::dec(tu, dp.invscale[1]), vr = C::dec(tv, dp.invscale[2]), wr = C::dec(tw, dp.invscale[3]);
    stb(base + 1 * fs, boff, tu); stb(base + 2 * fs, boff, tv); stb(base + 3 * fs, boff, tw);
    stb(base + 4 * fs, boff, C::enc(o.xx - ur * ur, dp.scale[4]));
    stb(base + 5 * fs, boff, C::enc(o.yy - vr * vr, dp.scale[5]));
    stb(base + 6 * fs, boff, C::enc(o.zz - wr * wr, dp.scale[6]));
    stb(base + 7 * fs, boff, C::enc(o.xy - ur * vr, dp.scale[7]));
    stb(base + 8 * fs, boff, C::enc(o.xz - ur * wr, dp.scale[8]));
    stb(base + 9 * fs, boff, C::enc(o.yz - vr * wr, dp.scale[9]));
  } else {
    stb(base + 1 * fs, boff, C::enc(o.u, 0.f)); stb(base + 2 * fs, boff, C::enc(o.v, 0.f));
    stb(base + 3 * fs, boff, C::enc(o.w, 0.f));
    stb(base + 4 * fs, boff, C::enc(o.xx, 0.f)); stb(base + 5 * fs, boff, C::enc(o.yy, 0.f));
    stb(base + 6 * fs, boff, C::enc(o.zz, 0.f)); stb(base + 7 * fs, boff, C::enc(o.xy, 0.f));
    stb(base + 8 * fs, boff, C::enc(o.xz, 0.f)); stb(base + 9 * fs, boff, C::enc(o.yz, 0.f));
  }
}

// Buffer-addressed variants: plane0 = field 0 of the set at the wanted z plane.
template <int P>
__device__ __forceinline__ Mom<typename Codec<P>::R> load_mom_b(const typename Codec<P>::T* plane0, size_t fs,
                                                                uint32_t pb, uint32_t boff, const DevParams& dp) {
  typedef Codec<P> C; typedef typename C::R R;
  Mom<R> o;
  o.P = C::dec(fld(plane0, fs, 0, pb, boff), dp.invscale[0]);
  o.u = C::dec(fld(plane0, fs, 1, pb, boff), dp.invscale[1]);
  o.v = C::dec(fld(plane0, fs, 2, pb, boff), dp.invscale[2]);
  o.w = C::dec(fld(plane0, fs, 3, pb, boff), dp.invscale[3]);
  o.xx = C::dec(fld(plane0, fs, 4, pb, boff), dp.invscale[4]);
  o.yy = C::dec(fld(plane0, fs, 5, pb, boff), dp.invscale[5]);
  o.zz = C::dec(fld(plane0, fs, 6, pb, boff), dp.invscale[6]);
  o.xy = C::dec(fld(plane0, fs, 7, pb, boff), dp.invscale[7]);
  o.xz = C::dec(fld(plane0, fs, 8, pb, boff), dp.invscale[8]);
  o.yz = C::dec(fld(plane0, fs, 9, pb, boff), dp.invscale[9]);
  if constexpr (C::dev) {
    o.xx = C::decpi(fld(plane0, fs, 4, pb, boff), dp.invscale[4], o.u * o.u);
    o.yy = C::decpi(fld(plane0, fs, 5, pb, boff), dp.invscale[5], o.v * o.v);
    o.zz = C::decpi(fld(plane0, fs, 6, pb, boff), dp.invscale[6], o.w * o.w);
    o.xy = C::decpi(fld(plane0, fs, 7, pb, boff), dp.invscale[7], o.u * o.v);
    o.xz = C::decpi(fld(plane0, fs, 8, pb, boff), dp.invscale[8], o.u * o.w);
    o.yz = C::decpi(fld(plane0, fs, 9, pb, boff), dp.invscale[9], o.v * o.w);
  }
  return o;
}

// Store of the ten moments; the 16-bit codecs round stochastically with the
// per-node, per-plane seed rnd.
template <int P>
__device__ __forceinline__ void store_mom_b(typename Codec<P>::T* plane0, size_t fs, uint32_t pb, uint32_t boff,
                                            const Mom<typename Codec<P>::R>& o, const DevParams& dp,
                                            uint32_t rnd) {
  typedef Codec<P> C; typedef typename C::R R; typedef typename C::T T;
  if constexpr (C::dev) {
    // each field's random bits: the FP32 bits of the next field of the same
    // group (cyclic), XOR the seed
    const uint32_t rm = rnd & C::RBITS;
    auto sr = [&](float x, float xn) { return C::srx(x, (__float_as_uint(xn) & C::RBITS) ^ rm); };
    const float x0 = o.P * scl<P>(dp, 0), x1 = o.u * scl<P>(dp, 1), x2 = o.v * scl<P>(dp, 2), x3 = o.w * scl<P>(dp, 3);
    fst(plane0, fs, 0, pb, boff, sr(x0, x1));
    const T tu = sr(x1, x2), tv = sr(x2, x3), tw = sr(x3, x0);
    const R ur = C::dec(tu, iscl<P>(dp, 1)), vr = C::dec(tv, iscl<P>(dp, 2)), wr = C::dec(tw, iscl<P>(dp, 3));
    fst(plane0, fs, 1, pb, boff, tu); fst(plane0, fs, 2, pb, boff, tv); fst(plane0, fs, 3, pb, boff, tw);
    const float x4 = (o.xx - ur * ur) * scl<P>(dp, 4), x5 = (o.yy - vr * vr) * scl<P>(dp, 5);
    const float x6 = (o.zz - wr * wr) * scl<P>(dp, 6), x7 = (o.xy - ur * vr) * scl<P>(dp, 7);
    const float x8 = (o.xz - ur * wr) * scl<P>(dp, 8), x9 = (o.yz - vr * wr) * scl<P>(dp, 9);
    fst(plane0, fs, 4, pb, boff, sr(x4, x5));
    fst(plane0, fs, 5, pb, boff, sr(x5, x6));
    fst(plane0, fs, 6, pb, boff, sr(x6, x7));
    fst(plane0, fs, 7, pb, boff, sr(x7, x8));
    fst(plane0, fs, 8, pb, boff, sr(x8, x9));
    fst(plane0, fs, 9, pb, boff, sr(x9, x4));
  } else {
    fst(plane0, fs, 0, pb, boff, C::enc(o.P, 0.f));
    fst(plane0, fs, 1, pb, boff, C::enc(o.u, 0.f)); fst(plane0, fs, 2, pb, boff, C::enc(o.v, 0.f));
    fst(plane0, fs, 3, pb, boff, C::enc(o.w, 0.f));
    fst(plane0, fs, 4, pb, boff, C::enc(o.xx, 0.f)); fst(plane0, fs, 5, pb, boff, C::enc(o.yy, 0.f));
    fst(plane0, fs, 6, pb, boff, C::enc(o.zz, 0.f)); fst(plane0, fs, 7, pb, boff, C::enc(o.xy, 0.f));
    fst(plane0, fs, 8, pb, boff, C::enc(o.xz, 0.f)); fst(plane0, fs, 9, pb, boff, C::enc(o.yz, 0.f));
  }
}

// ---------------------------------------------------------------------------
// Lattice.  Populations are addressed by velocity (p,q,r) in {-1,0,1}^3
// through cube index c = (p+1) + 3(q+1) + 9(r+1); the reference's ordering
// (vars_module.f90:319-338, D3Q19 = first 19 entries) matters only for
// loops that follow its order, which use their own tables.
__host__ __device__ constexpr int cidx(int p, int q, int r) { return (p + 1) + 3 * (q + 1) + 9 * (r + 1); }

template <int LAT>
__host__ __device__ constexpr bool present(int p, int q, int r) {
  // D3Q27: all 27; D3Q19: drop the 8 corners
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
// The centre value is formed as A0 - (fm + fp), which equals 2/3 A0 - A2:
// in FP32, fl(1/6) + fl(2/3) + fl(1/6) = 1 + 3.0e-8, so weighting A0 by the
// rounded 2/3 would scale the zeroth moment (the density deviation P) by
// 1 + 9e-8 per step over the three passes -- a mass drift eleven times that
// of the rounded weights of the reference's own table.  Formed this way the
// three values sum to A0 up to unbiased rounding.
template <class R>
__device__ __forceinline__ void recon3(R A0, R A1, R A2, R& fm, R& f0, R& fp) {
  const R t = A0 * (R)(1.0 / 6.0) + A2 * (R)0.5;
  fp = t + A1 * (R)0.5;
  fm = t - A1 * (R)0.5;
  f0 = A0 - (R)2 * t;
}

}  // namespace lbf
