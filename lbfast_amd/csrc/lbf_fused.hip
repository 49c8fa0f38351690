// lbf_fused.hip -- fused stream+collide from moments for gfx950.
//
// Replaces fused_LB_kernel (lb_cuda_fused_module.f90:426-750) and, in the
// "inline" variant used by lbf_step, also the moments stage
// (lb_cuda_moments_module.f90:19-341), the phase-field update
// (lb_cuda_update_phi_module.f90:19-104) and the near-wall regularised
// bounce-back (lb_cuda_boundary_module.f90:19-216), so that one launch reads
// the state once and writes the new state once.
//
// Arithmetic.  The reference rebuilds 27 post-collision populations as
//   f*_l = feq_l + (1-omega) fneq_l + F_l/2
// with feq/fneq the second-order terms plus, under HIGHORDER, Minv_mrt(k,l)
// times the monomials E_k and polynomials N_k (k = 10..26, fused:641-702).
// Every term is a Hermite coefficient times Minv_mrt(k,l) = w_l H_k(c_l)/b_k,
// and for D3Q27 the basis is the tensor product of the 1-D basis
// {1, c, c^2-1/3} (vars_module.f90:522-664), so
//   f*_{pqr} = sum_{abc} g_a(p) g_b(q) g_c(r) A_{abc},
//   A_0 = P, A_i = u_i + a_i/2, A_ij = u_i u_j + (1-omega) Pn_ij + (u_i a_j + u_j a_i)/2,
//   A_k = E_k + (1-omega) N_k (k >= 10)
// (a = force/rho_phi; Guo forcing of fused:705-706 is exactly the order-1/2
// Hermite part).  The 27 populations then cost three 1-D passes (135 ops)
// instead of the dense 27x27 sum.  D3Q19 is not a tensor product; its 19
// populations are built directly from the order <= 2 coefficients.
//
// Streaming + moments.  out_m(x) = sum_l M_ml f*_l(x - c_l) separates by
// axis too.  A block owns a 64-wide x segment (one wave per y row, lanes =
// x) of TY output rows plus one halo row on each side, and marches z.  For
// each source plane it:
//   x: shifts the p = +-1 populations one lane (DPP) and contracts
//      over p with (1, c, c^2-1/3); the two x-halo nodes of every row are
//      computed by one extra wave and handed over through LDS;
//   y: exchanges the q = +-1 partial sums with the neighbouring rows through
//      LDS and contracts over q;
//   z: adds the r = -1/0/+1 partial sums into three per-lane accumulator
//      planes; the lowest is complete and is stored.
// Every node of the block's (x,y) footprint is collided once per plane; the
// recomputation is only the two y-halo rows, the x-halo wave and the two
// z-halo planes of the march.
//
// Phase-field inputs (two-component, inline variant).  Without internal
// obstacles and with float phase-field storage the block also streams a phi
// tile with a 2-node halo and forms grad phi and lap phi of every collided
// node from the separable partial sums of three planes (part3 / grad_lap,
// the arithmetic of the prep kernel), so the prep only stores div theta.
// With obstacles, PHI_int_boundary (lb_cuda_boundary_module.f90:218-459)
// rewrites phi on solid nodes after compute_norm and before the moments
// stage, so the gradients come from the prep's output instead.
//
// Periodicity: indices wrap in-kernel (the reference's intpbc halo copies,
// mpi_module.f90:2820-2918).  isfluid is read from a copy with the
// reference's 1-node halo (periodic image or 1, LBFAST.f90:258-328).
#include "lbf_common.h"
#include "lbf_kernels.h"
#include "lbf_physics.h"

namespace lbf {

// Whole-wave lane shifts through DPP (wave_shr:1 / wave_shl:1, a VALU
// operand modifier, no LDS round trip).  With bound_ctrl off, the lane whose
// source lies outside the wave (lane 0 for shr, lane 63 for shl) keeps `fill`.
__device__ __forceinline__ int dpp_shr1(int fill, int v) { return __builtin_amdgcn_update_dpp(fill, v, 0x138, 0xF, 0xF, false); }
__device__ __forceinline__ int dpp_shl1(int fill, int v) { return __builtin_amdgcn_update_dpp(fill, v, 0x130, 0xF, 0xF, false); }
__device__ __forceinline__ float shift_from_lower(float v, float fill) {  // lane L receives lane L-1
  return __builtin_bit_cast(float, dpp_shr1(__builtin_bit_cast(int, fill), __builtin_bit_cast(int, v)));
}
__device__ __forceinline__ float shift_from_upper(float v, float fill) {  // lane L receives lane L+1
  return __builtin_bit_cast(float, dpp_shl1(__builtin_bit_cast(int, fill), __builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double shift_from_lower(double v, double fill) {
  const int2 t = __builtin_bit_cast(int2, v), u = __builtin_bit_cast(int2, fill);
  return __builtin_bit_cast(double, make_int2(dpp_shr1(u.x, t.x), dpp_shr1(u.y, t.y)));
}
__device__ __forceinline__ double shift_from_upper(double v, double fill) {
  const int2 t = __builtin_bit_cast(int2, v), u = __builtin_bit_cast(int2, fill);
  return __builtin_bit_cast(double, make_int2(dpp_shl1(u.x, t.x), dpp_shl1(u.y, t.y)));
}
// Encoded inputs of one node for one source plane.
template <int PREC>
struct RawNode {
  typename Codec<PREC>::T m[10];
  typename Codec<PREC>::T uo[3];
  typename Aux<PREC>::T phi;
  typename Aux<PREC>::T g[5];  // inline: grad phi (3), lap phi, div theta n; stage path: forces (3)
};

// LDS slots of the prefetched inputs (one 64-lane row of dwords each):
// moments 0-9, phi 10, g 11-15, u_old 16-18.  GI (gradient inline): phi
// comes from the block's phi tile instead, and g holds div theta only (slot
// 10, u_old 11-13).
template <bool TWO, bool INLINE, bool GI = false>
struct Slots {
  static constexpr int PHI = 10, G = 11, UO = 16;
  static constexpr int N = INLINE ? (TWO ? 19 : 10) : 14;
};
template <>
struct Slots<true, true, true> {
  static constexpr int PHI = -1, G = 10, UO = 11, N = 14;
};

// Per-lane byte offsets of one node inside a z plane of each input buffer,
// and the planes' sizes in bytes: the 10-field moment set, the 3-field
// u_old set, the g set (prep output, 5 fields, or forces, 3 fields) and the
// single phase field.  Field sets are row-interleaved, so a set's field f is
// the same lane offset from a plane base shifted by f*sx elements.
struct FOff {
  uint32_t m, u, g, p;
  uint32_t pm, pu, pg, pp;
  // 16-bit storage, full 64-node row segments (pair != 0, wave-uniform;
  // lane l holds node x0 + l): one transfer moves two fields' rows, lanes
  // 0-31 the dwords of field f (nodes x0 .. x0+63) and lanes 32-63 those of
  // field f+1 (pair_delta)
  int pair;
};

// byte offset from a lane's own 16-bit element to its dword of a paired
// transfer: lane l < 32 fetches elements x0+2l, x0+2l+1 of field f, lane
// l >= 32 elements x0+2(l-32), +1 of field f+1 (sx elements further)
__device__ __forceinline__ uint32_t pair_delta(int sx) {
  const uint32_t l = threadIdx.x & 63;
  return 2u * (l < 32 ? l : l - 64u + (uint32_t)sx);
}

// Direct loads of one plane's inputs (used when the inputs are not prefetched).
// fs = sx, the field stride of the sets.
template <int PREC, bool TWO, bool INLINE>
__device__ __forceinline__ void load_raw(RawNode<PREC>& r, const FusedArgs& a, int zw, size_t fs, const FOff& o) {
  typedef typename Codec<PREC>::T T;
  typedef typename Aux<PREC>::T TA;
  constexpr int GF = INLINE ? 5 : 3;
  const DevParams& dp = a.dp;
  const size_t rows = (size_t)zw * dp.ny;
  const T* S = (const T*)a.min + rows * 10 * fs;
#pragma unroll
  for (int f = 0; f < 10; ++f) r.m[f] = fld(S, fs, f, o.pm, o.m);
  if constexpr (TWO) r.phi = fld((const TA*)a.phin + rows * dp.nx, fs, 0, o.pp, o.p);
  if constexpr (!INLINE) {
    const TA* fr = (const TA*)a.frc + rows * GF * fs;
#pragma unroll
    for (int f = 0; f < 3; ++f) r.g[f] = fld(fr, fs, f, o.pg, o.g);
  } else if constexpr (TWO) {
    const TA* pr = (const TA*)a.pre + rows * GF * fs;
#pragma unroll
    for (int f = 0; f < 5; ++f) r.g[f] = fld(pr, fs, f, o.pg, o.g);
    if (dp.dens) {
      const T* uo = (const T*)a.uin + rows * 3 * fs;
#pragma unroll
      for (int f = 0; f < 3; ++f) r.uo[f] = fld(uo, fs, f, o.pu, o.u);
    }
  }
}

// LDS-DMA prefetch of one plane's inputs into this wave's slots.  Float
// fields move their lane's dword; a 16-bit moment field moves the aligned
// dword that holds the lane's element (the half is picked when read back:
// byte (o.m & 2) of the slot; sx is even, so every 16-bit field row starts on
// a dword and the u_old offsets share that parity).
template <int PREC, bool TWO, bool INLINE, bool GI>
__device__ __forceinline__ void dma_raw(const FusedArgs& a, int zw, size_t fs, const FOff& o, uint32_t lds0) {
  typedef Slots<TWO, INLINE, GI> SL;
  typedef typename Codec<PREC>::T T;
  static_assert(sizeof(typename Aux<PREC>::T) == 4, "float phase-field storage");
  constexpr int GF = INLINE ? 5 : 3;
  const DevParams& dp = a.dp;
  const size_t rows = (size_t)zw * dp.ny;
  const uint32_t vM = o.m & ~3u, vU = o.u & ~3u;
  const T* S = (const T*)a.min + rows * 10 * fs;
  // slot offsets are immediates: only the wave's LDS base stays live
  constexpr bool P16 = sizeof(T) == 2 && INLINE;
  if (P16 && o.pair) {
    // field pairs (0,1) .. (8,9) into slots 0-4
    const uint32_t vP = o.m + pair_delta(dp.sx);
    dma_dword<256u * 0>(mkrsrc(S + 0 * fs, o.pm), vP, lds0);
    dma_dword<256u * 1>(mkrsrc(S + 2 * fs, o.pm), vP, lds0);
    dma_dword<256u * 2>(mkrsrc(S + 4 * fs, o.pm), vP, lds0);
    dma_dword<256u * 3>(mkrsrc(S + 6 * fs, o.pm), vP, lds0);
    dma_dword<256u * 4>(mkrsrc(S + 8 * fs, o.pm), vP, lds0);
  } else {
    dma_dword<256u * 0>(mkrsrc(S + 0 * fs, o.pm), vM, lds0);
    dma_dword<256u * 1>(mkrsrc(S + 1 * fs, o.pm), vM, lds0);
    dma_dword<256u * 2>(mkrsrc(S + 2 * fs, o.pm), vM, lds0);
    dma_dword<256u * 3>(mkrsrc(S + 3 * fs, o.pm), vM, lds0);
    dma_dword<256u * 4>(mkrsrc(S + 4 * fs, o.pm), vM, lds0);
    dma_dword<256u * 5>(mkrsrc(S + 5 * fs, o.pm), vM, lds0);
    dma_dword<256u * 6>(mkrsrc(S + 6 * fs, o.pm), vM, lds0);
    dma_dword<256u * 7>(mkrsrc(S + 7 * fs, o.pm), vM, lds0);
    dma_dword<256u * 8>(mkrsrc(S + 8 * fs, o.pm), vM, lds0);
    dma_dword<256u * 9>(mkrsrc(S + 9 * fs, o.pm), vM, lds0);
  }
  if constexpr (TWO && !GI) dma_dword<256u * SL::PHI>(mkrsrc((const float*)a.phin + rows * dp.nx, o.pp), o.p, lds0);
  if constexpr (!INLINE) {
    const float* fr = (const float*)a.frc + rows * GF * fs;
    dma_dword<256u * (SL::G + 0)>(mkrsrc(fr + 0 * fs, o.pg), o.g, lds0);
    dma_dword<256u * (SL::G + 1)>(mkrsrc(fr + 1 * fs, o.pg), o.g, lds0);
    dma_dword<256u * (SL::G + 2)>(mkrsrc(fr + 2 * fs, o.pg), o.g, lds0);
  } else if constexpr (TWO) {
    const float* pr = (const float*)a.pre + rows * GF * fs;
    if constexpr (GI) {
      dma_dword<256u * SL::G>(mkrsrc(pr + 4 * fs, o.pg), o.g, lds0);
    } else {
      dma_dword<256u * (SL::G + 0)>(mkrsrc(pr + 0 * fs, o.pg), o.g, lds0);
      dma_dword<256u * (SL::G + 1)>(mkrsrc(pr + 1 * fs, o.pg), o.g, lds0);
      dma_dword<256u * (SL::G + 2)>(mkrsrc(pr + 2 * fs, o.pg), o.g, lds0);
      dma_dword<256u * (SL::G + 3)>(mkrsrc(pr + 3 * fs, o.pg), o.g, lds0);
      dma_dword<256u * (SL::G + 4)>(mkrsrc(pr + 4 * fs, o.pg), o.g, lds0);
    }
    if (dp.dens) {
      const T* uo = (const T*)a.uin + rows * 3 * fs;
      if (P16 && o.pair) {
        // (u, v) into slot UO, w into the lower half of slot UO+1
        const uint32_t vP = o.u + pair_delta(dp.sx);
        dma_dword<256u * (SL::UO + 0)>(mkrsrc(uo + 0 * fs, o.pu), vP, lds0);
        if ((threadIdx.x & 63) < 32) dma_dword<256u * (SL::UO + 1)>(mkrsrc(uo + 2 * fs, o.pu), vP, lds0);
      } else {
        dma_dword<256u * (SL::UO + 0)>(mkrsrc(uo + 0 * fs, o.pu), vU, lds0);
        dma_dword<256u * (SL::UO + 1)>(mkrsrc(uo + 1 * fs, o.pu), vU, lds0);
        dma_dword<256u * (SL::UO + 2)>(mkrsrc(uo + 2 * fs, o.pu), vU, lds0);
      }
    }
  }
}

template <int PREC, bool TWO, bool INLINE, bool GI>
__device__ __forceinline__ void lds_raw(RawNode<PREC>& r, const float (*slot)[64], int lane, bool dens, uint32_t bT,
                                        bool pair) {
  typedef Slots<TWO, INLINE, GI> SL;
  typedef typename Codec<PREC>::T T;
  constexpr bool P16 = sizeof(T) == 2 && INLINE;
  // a 16-bit element sits at byte (bT & 2) of its lane's dword; in a paired
  // slot (s0 + f/2) field f's 64 elements are the slot's lower or upper 128 B
  const uint32_t hs = sizeof(T) == 2 ? (bT & 2u) : 0u;
  auto one = [&](int f) { return *(const T*)((const char*)&slot[f][lane] + hs); };
  auto two = [&](int s0, int f) { return *(const T*)((const char*)&slot[s0 + (f >> 1)][0] + (f & 1) * 128 + lane * 2); };
  const bool pr = P16 && pair;
  constexpr bool UO = INLINE && TWO;
  if (pr) {
#pragma unroll
    for (int f = 0; f < 10; ++f) r.m[f] = two(0, f);
    if (UO && dens) {
#pragma unroll
      for (int f = 0; f < 3; ++f) r.uo[f] = two(SL::UO, f);
    }
  } else {
#pragma unroll
    for (int f = 0; f < 10; ++f) r.m[f] = one(f);
    if (UO && dens) {
#pragma unroll
      for (int f = 0; f < 3; ++f) r.uo[f] = one(SL::UO + f);
    }
  }
  if constexpr (TWO && !GI) r.phi = slot[SL::PHI][lane];
  if constexpr (!INLINE) {
#pragma unroll
    for (int f = 0; f < 3; ++f) r.g[f] = slot[SL::G + f][lane];
  } else if constexpr (GI) {
    r.g[4] = slot[SL::G][lane];
  } else if constexpr (TWO) {
#pragma unroll
    for (int f = 0; f < 5; ++f) r.g[f] = slot[SL::G + f][lane];
  }
}

// ---------------------------------------------------------------------------
// PF: the next plane's inputs stream into LDS (dma_raw) while the current
// plane is collided and streamed.  All stores of an iteration are issued
// before that DMA, so the wait at the top of the next iteration (which the
// shared load/store counter makes a wait for everything) finds the older
// stores long done.  The completed output plane is therefore stored at the
// top of the following iteration.
template <int TY>
__host__ __device__ inline int fused_tiles(const FusedArgs& a) {
  return ((a.dp.nx + 63) / 64) * ((a.dp.ny + TY - 1) / TY) * ((a.dp.nz + a.tz - 1) / a.tz);
}

// One block: tile `tile` (x fastest, then y, then z chunks) of the domain `a`.
// OB: the domain has internal obstacles (isfluid loads, solid nodes, the
// near-wall closures); without them that code is compiled out.
// A node next to the wall of a non-periodic axis (coordinate 1 or n-2): a
// superset of the first fluid layers of the open faces.
__device__ __forceinline__ bool open_layer(const DevParams& dp, int x, int y, int z) {
  return (!dp.pbc[0] && (x == 1 || x == dp.nx - 2)) || (!dp.pbc[1] && (y == 1 || y == dp.ny - 2)) ||
         (!dp.pbc[2] && (z == 1 || z == dp.nz - 2));
}

// OPEN: open faces (with OB and the inline moments stage only): the kernel
// also hands over the post-moments state of the wall-adjacent nodes
template <int LAT, bool HO, int PREC, bool TWO, bool INLINE, int TY, bool OB, bool OPEN = false>
__device__ __forceinline__ void fused_body(const FusedArgs& a, int tile) {
  typedef Codec<PREC> C;
  typedef typename C::R R;
  typedef typename C::T T;
  typedef typename Aux<PREC>::T TA;
  constexpr int NW = TY + 2;  // row waves (TY output rows + 2 halo rows)
  constexpr int PB = prec_base(PREC);
  constexpr bool PF = PB == PREC_FP32 || PB == PREC_FP16S || PB == PREC_BF16S;
  // GI: grad phi and lap phi formed here from a phi tile (fused_grad_inline)
  constexpr bool GI = PF && TWO && INLINE && !OB;
  typedef Slots<TWO, INLINE, GI> SL;
  // GI phi tile: rows y0-2 .. y0+TY+1 (tile row r holds y0-2+r), columns
  // x0-2 .. x0+65 (column c holds x0-2+c), both wrapped; a ring of four z
  // planes streamed by LDS DMA three planes ahead of the plane collided
  constexpr int GR = TY + 4, GC = 68;

  __shared__ R sUP[NW][9][64];
  __shared__ R sDN[NW][9][64];
  __shared__ R sHX[NW][2][9];
  __shared__ float sPF[PF ? NW + 1 : 1][PF ? SL::N : 1][64];
  __shared__ float sPH[GI ? 4 : 1][GI ? GR : 1][GI ? GC : 1];

  const DevParams& dp = a.dp;
  const Phys<R>& ph = dp.ph<R>();
  const int nx = dp.nx, ny = dp.ny, nz = dp.nz;
  const size_t nxny = (size_t)nx * ny;
  const bool obst = OB && dp.obst != 0, bb = OB && dp.bb != 0;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int ntx = (nx + 63) / 64, nty = (ny + TY - 1) / TY;
  const int tx = tile % ntx, ty = (tile / ntx) % nty, tzi = tile / (ntx * nty);
  const int x0 = tx * 64;
  const int XW = min(64, nx - x0);
  const int y0 = ty * TY;
  const int z0 = tzi * a.tz;
  const int zend = min(z0 + a.tz, nz);
  const bool halo_wave = (wave == NW);

  int w, xs, side = 0;
  bool active = true;
  if (!halo_wave) {
    w = wave;
    xs = x0 + lane;
  } else {
    w = lane >> 1;
    side = lane & 1;
    active = lane < 2 * NW;
    if (!active) w = 0;
    xs = side ? x0 + XW : x0 - 1;
  }
  const int ys = y0 - 1 + w;
  // row waves: w == wave (wave-uniform); the halo wave's lanes map to (row, side)
  const bool out_row = !halo_wave && wave >= 1 && wave <= TY && y0 - 1 + wave < ny;
  const bool out_node = out_row && lane < XW;
  const int xw = ((xs % nx) + nx) % nx;  // lanes past a short row wrap more than once
  const int yw = ((ys % ny) + ny) % ny;  // ys may exceed 2ny when ny < TY
  constexpr int GF = INLINE ? 5 : 3;     // fields of the g set: prep output or forces
  const int sx = dp.sx;
  const uint32_t srow = (uint32_t)sx * (uint32_t)yw;
  FOff off;
  off.m = ((uint32_t)xw + 10u * srow) * (uint32_t)sizeof(T);
  off.u = ((uint32_t)xw + 3u * srow) * (uint32_t)sizeof(T);
  // paired 16-bit transfers for row waves over a full 64-node segment
  // (x0 + 63 < nx: no wrap inside the segment)
  off.pair = (!halo_wave && XW == 64) ? 1 : 0;
  off.g = ((uint32_t)xw + (uint32_t)GF * srow) * (uint32_t)sizeof(TA);
  off.p = ((uint32_t)xw + (uint32_t)nx * (uint32_t)yw) * (uint32_t)sizeof(TA);
  const uint32_t splane = (uint32_t)sx * (uint32_t)ny;
  off.pm = ((10u * splane * (uint32_t)sizeof(T)) + 3u) & ~3u;
  off.pu = ((3u * splane * (uint32_t)sizeof(T)) + 3u) & ~3u;
  off.pg = (uint32_t)GF * splane * (uint32_t)sizeof(TA);
  off.pp = (uint32_t)(nxny * sizeof(TA));
  // isfluid with halo: per-lane in-plane offset, per-plane uniform base
  const uint32_t bI = (uint32_t)(clampi(xs, -1, nx) + 1) + (uint32_t)(nx + 2) * (uint32_t)(clampi(ys, -1, ny) + 1);
  const size_t hplane = (size_t)(nx + 2) * (ny + 2);
  const uint32_t pbI = (uint32_t)hplane;
  auto isf_load = [&](int z) -> int {
    return (int)bld<int8_t>(mkrsrc(a.isfh + (size_t)(clampi(z, -1, nz) + 1) * hplane, pbI), bI);
  };

  T* Sout = (T*)a.mout;
  const size_t plM = (size_t)10 * sx * ny;  // elements of one z plane of the moment set
  // 16-bit codecs: per-node seed of the stochastic rounding
  // and of the step: without a step term a node whose sibling field is
  // constant (e.g. exactly 0 in a planar flow) would draw the same threshold
  // every step, which makes its rounding deterministic and biased
  uint32_t nseed = 0u;
  if constexpr (C::dev) {
    const uint32_t step = a.seed + (a.seedp ? *a.seedp : 0u);
    nseed = mix32((uint32_t)xw * 0x9E3779B9u ^ (uint32_t)(yw + 1) * 0x85EBCA6Bu ^ mix32(step * 0x632BE5ABu + 0x1B873593u));
  }

  R lo[10], mid[10], hi[10];
#pragma unroll
  for (int m = 0; m < 10; ++m) { lo[m] = 0; mid[m] = 0; hi[m] = 0; }
  auto store_lo = [&](int zo, size_t fs) {
    Mom<R> o;
    o.P = lo[0]; o.u = lo[1]; o.v = lo[2]; o.w = lo[3];
    o.xx = lo[4]; o.yy = lo[5]; o.zz = lo[6]; o.xy = lo[7]; o.xz = lo[8]; o.yz = lo[9];
    store_mom_b<PREC>(Sout + (size_t)zo * plM, fs, off.pm, off.m, o, dp, nseed ^ (uint32_t)zo * 0xC2B2AE35u);
  };

  const R third = (R)(1.0 / 3.0), twothird = (R)(2.0 / 3.0);

  uint32_t lds0 = 0;
  int isf_next = 1;
  if constexpr (PF) {
    lds0 = lds_addr(&sPF[wave][0][0]);
    dma_raw<PREC, TWO, INLINE, GI>(a, wrapg(z0 - 1, nz), (size_t)sx, off, lds0);
    if (obst) isf_next = isf_load(z0 - 1);
  }

  // ---- GI: phi tile stream and the per-node partial sums ------------------------
  // Wave v streams tile row v (v = 0 .. TY+2; the x-halo wave is v = NW), wave
  // 0 also row TY+3: 64 lanes to columns 2..65 and lanes 0-1 to columns 0-1
  // and 66-67.  Plane p lands in slot p & 3.  Iteration zsrc reads plane
  // zsrc+1 (issued in iteration zsrc-2, waited for at the top of zsrc-1 and
  // published by that iteration's barriers) and streams plane zsrc+3 into the
  // slot of plane zsrc-1, which every wave finished reading in iteration
  // zsrc-2.
  uint32_t ph_row1 = 0, ph_row2 = 0, ph_xm = 0, ph_xl = 0, ph_xr = 0;
  const uint32_t lds_ph = GI ? lds_addr(&sPH[0][0][0]) : 0u;
  auto modn = [](int i, int n) { return ((i % n) + n) % n; };
  if constexpr (GI) {
    ph_row1 = __builtin_amdgcn_readfirstlane((uint32_t)modn(y0 - 2 + wave, ny) * (uint32_t)nx * 4u);
    ph_row2 = __builtin_amdgcn_readfirstlane((uint32_t)modn(y0 - 2 + GR - 1, ny) * (uint32_t)nx * 4u);
    ph_xm = (uint32_t)modn(x0 + lane, nx) * 4u;
    ph_xl = (uint32_t)modn(x0 - 2 + (lane & 1), nx) * 4u;
    ph_xr = (uint32_t)modn(x0 + 64 + (lane & 1), nx) * 4u;
  }
  auto dma_phi = [&](int z) {
    const Rsrc rs = mkrsrc((const float*)a.phin + (size_t)modn(z, nz) * nxny, (uint32_t)(nxny * 4u));
    const uint32_t base = lds_ph + (uint32_t)(z & 3) * (uint32_t)(GR * GC * 4);
    const uint32_t r1 = base + (uint32_t)wave * (uint32_t)(GC * 4);
    dma_dword<8u>(rs, ph_row1 + ph_xm, r1);
    if (lane < 2) {
      dma_dword<0u>(rs, ph_row1 + ph_xl, r1);
      dma_dword<66u * 4u>(rs, ph_row1 + ph_xr, r1);
    }
    if (wave == 0) {
      const uint32_t r2 = base + (uint32_t)(GR - 1) * (uint32_t)(GC * 4);
      dma_dword<8u>(rs, ph_row2 + ph_xm, r2);
      if (lane < 2) {
        dma_dword<0u>(rs, ph_row2 + ph_xl, r2);
        dma_dword<66u * 4u>(rs, ph_row2 + ph_xr, r2);
      }
    }
  };
  // this lane's node sits at tile row w + 1, column pc
  const int pc = halo_wave ? (side ? XW + 2 : 1) : lane + 2;
  auto phi_part = [&](int z) -> Part3<R> {
    const int sl = z & 3;
    R v[3][3];
#pragma unroll
    for (int bb = 0; bb < 3; ++bb)
#pragma unroll
      for (int aa = 0; aa < 3; ++aa) v[bb][aa] = (R)sPH[sl][w + bb][pc - 1 + aa];
    return part3(v);
  };
  Part3<R> gpm{}, gp0{};  // partials of planes zsrc-1 and zsrc (gp0.c: phi at the node)
  if constexpr (GI) {
    dma_phi(z0 - 2);
    dma_phi(z0 - 1);
    dma_phi(z0);
    dma_phi(z0 + 1);
    dma_wait_all();
    __syncthreads();
    gpm = phi_part(z0 - 2);
    gp0 = phi_part(z0 - 1);
    __syncthreads();  // the slots of planes z0-2 and z0-1 are refilled in the first iterations
  }

  for (int zsrc = z0 - 1; zsrc <= zend; ++zsrc) {
    const int zw = wrapg(zsrc, nz);
    const bool own_plane = zsrc >= z0 && zsrc < zend;
    // the field stride, opaque to the optimiser inside the loop: the ~30
    // per-field plane offsets are then rebuilt by scalar ALU each iteration
    // instead of being hoisted and spilled to vector-register lanes
    size_t fs = (size_t)sx;
    asm volatile("" : "+s"(fs));

    RawNode<PREC> cur;
    int8_t isf = 1;
    if constexpr (PF) {
      dma_wait_all();
      // 16-bit storage: waves between this wait and the next plane's DMA
      // issue (store, decode, moments stage) run at raised priority, ahead of
      // waves in the collision, so the transfers start sooner (fused kernel
      // 3.92 -> 3.43 ms FP16S at 512^3; no gain for FP32)
      if constexpr (sizeof(T) == 2) __builtin_amdgcn_s_setprio(2);
      int t = isf_next;
      asm volatile("" : "+v"(t));  // consume the prefetched isfluid before this iteration's stores
      isf = (int8_t)t;
      // the plane completed by the previous iteration
      if (out_row && lane < XW && zsrc - 2 >= z0) store_lo(zsrc - 2, fs);
#pragma unroll
      for (int m = 0; m < 10; ++m) { lo[m] = mid[m]; mid[m] = hi[m]; }
      lds_raw<PREC, TWO, INLINE, GI>(cur, sPF[wave], lane, dp.dens, off.m, off.pair != 0);
      if constexpr (GI) {
        // the prep's fields 0-3 (prep_body_z, stored at storage precision)
        const Part3<R> gpp = phi_part(zsrc + 1);
        R gx, gy, gz, lap;
        grad_lap(gpm, gp0, gpp, gx, gy, gz, lap);
        const R mg = sqrt(gx * gx + gy * gy + gz * gz);
        const R im = rcp(mg + (R)1.0e-9);
        const R mgs = (R)(TA)mg;
        cur.g[0] = (TA)((R)(TA)(gx * im) * mgs);
        cur.g[1] = (TA)((R)(TA)(gy * im) * mgs);
        cur.g[2] = (TA)((R)(TA)(gz * im) * mgs);
        cur.g[3] = (TA)lap;
        cur.phi = (TA)gp0.c;
        gpm = gp0;
        gp0 = gpp;
      }
    } else {
      if (obst) isf = (int8_t)isf_load(zsrc);
      load_raw<PREC, TWO, INLINE>(cur, a, zw, fs, off);
    }

    // ---- 1. decode + (inline) moments stage + collision ---------------------
    R f[27];
    R lbc[10];
    bool has_lbc = false;
    {
      NodeIn<R> in;
      Mom<R> m;
      m.P = C::dec(cur.m[0], iscl<PREC>(dp, 0)); m.u = C::dec(cur.m[1], iscl<PREC>(dp, 1));
      m.v = C::dec(cur.m[2], iscl<PREC>(dp, 2)); m.w = C::dec(cur.m[3], iscl<PREC>(dp, 3));
      m.xx = C::dec(cur.m[4], iscl<PREC>(dp, 4)); m.yy = C::dec(cur.m[5], iscl<PREC>(dp, 5));
      m.zz = C::dec(cur.m[6], iscl<PREC>(dp, 6)); m.xy = C::dec(cur.m[7], iscl<PREC>(dp, 7));
      m.xz = C::dec(cur.m[8], iscl<PREC>(dp, 8)); m.yz = C::dec(cur.m[9], iscl<PREC>(dp, 9));
      if constexpr (C::dev) {
        // second moments: stored deviation + u u (one mixed FMA each)
        m.xx = C::decpi(cur.m[4], iscl<PREC>(dp, 4), m.u * m.u);
        m.yy = C::decpi(cur.m[5], iscl<PREC>(dp, 5), m.v * m.v);
        m.zz = C::decpi(cur.m[6], iscl<PREC>(dp, 6), m.w * m.w);
        m.xy = C::decpi(cur.m[7], iscl<PREC>(dp, 7), m.u * m.v);
        m.xz = C::decpi(cur.m[8], iscl<PREC>(dp, 8), m.u * m.w);
        m.yz = C::decpi(cur.m[9], iscl<PREC>(dp, 9), m.v * m.w);
      }
      R phl = 0, Fx, Fy, Fz;
      const bool solid = obst && isf == 0;
      if constexpr (TWO) phl = (R)cur.phi;
      const NodeRates<R> nr = node_rates<TWO, R>(dp, phl);
      const R rinv = nr.irho;
      if constexpr (!INLINE) {
        Fx = (R)cur.g[0] * rinv;
        Fy = (R)cur.g[1] * rinv;
        Fz = (R)cur.g[2] * rinv;
      } else {
        R gfx = 0, gfy = 0, gfz = 0, lap = 0, dvt = 0, uox = 0, uoy = 0, uoz = 0;
        if constexpr (TWO) {
          gfx = (R)cur.g[0];
          gfy = (R)cur.g[1];
          gfz = (R)cur.g[2];
          lap = (R)cur.g[3];
          dvt = (R)cur.g[4];
          if (dp.dens) {
            uox = C::dec(cur.uo[0], iscl<PREC>(dp, 1));
            uoy = C::dec(cur.uo[1], iscl<PREC>(dp, 2));
            uoz = C::dec(cur.uo[2], iscl<PREC>(dp, 3));
          }
        }
        R tFx, tFy, tFz;
        moments_stage<TWO, R>(dp, nr, phl, gfx, gfy, gfz, lap, dvt, uox, uoy, uoz, m, tFx, tFy, tFz);
        // the reference stores S and the forces at storage precision between
        // the moments and fused kernels
        m.u = C::round(m.u); m.v = C::round(m.v); m.w = C::round(m.w);
        m.xx = C::round(m.xx); m.yy = C::round(m.yy); m.zz = C::round(m.zz);
        m.xy = C::round(m.xy); m.xz = C::round(m.xz); m.yz = C::round(m.yz);
        tFx = C::round(tFx); tFy = C::round(tFy); tFz = C::round(tFz);
        Fx = tFx * rinv; Fy = tFy * rinv; Fz = tFz * rinv;
        if (out_node && own_plane && !solid) {
          if constexpr (OPEN) {
            if (isf == -1 && open_layer(dp, xs, ys, zsrc)) {
              // open faces: the post-moments state of a first-layer node,
              // which the open-face correction reads after this launch (the
              // stage sequence reads it from S)
              store_mom_b<PREC>((T*)a.sopen + (size_t)zw * plM, fs, off.pm, off.m, m, dp, 0u);
            }
          }
          if constexpr (TWO) {
            if (dp.dens) {
              T* uo = (T*)a.uout + (size_t)zw * ny * 3 * fs;
              T eu = C::enc(m.u, scl<PREC>(dp, 1)), ev = C::enc(m.v, scl<PREC>(dp, 2)), ew = C::enc(m.w, scl<PREC>(dp, 3));
              if constexpr (C::dev) {
                const uint32_t st = (nseed ^ (uint32_t)zw * 0x27D4EB2Fu) & C::RBITS;
                const float xu = m.u * scl<PREC>(dp, 1), xv = m.v * scl<PREC>(dp, 2), xw_ = m.w * scl<PREC>(dp, 3);
                eu = C::srx(xu, (__float_as_uint(xv) & C::RBITS) ^ st);
                ev = C::srx(xv, (__float_as_uint(xw_) & C::RBITS) ^ st);
                ew = C::srx(xw_, (__float_as_uint(xu) & C::RBITS) ^ st);
              }
              fst(uo, fs, 0, off.pu, off.u, eu);
              fst(uo, fs, 1, off.pu, off.u, ev);
              fst(uo, fs, 2, off.pu, off.u, ew);
            }
            // update_phi (update_phi:88-93), u after the moments stage
            const R t = -ph.sharp_c * dvt;
            const R po = phl - m.u * (R)0.5 * (gfx) - m.v * (R)0.5 * (gfy) - m.w * (R)0.5 * (gfz) +
                         ph.tau_diff * lap + t;
            fst((TA*)a.phiout + (size_t)zw * nxny, fs, 0, off.pp, off.p, (TA)po);
          }
          if (obst && !bb && isf == -1) {
            // LB_int_boundary (boundary:117-212) on the post-moments state
            has_lbc = true;
            const R cssq = (R)(1.0 / 3.0);
            const R pxx = m.xx - m.u * m.u, pyy = m.yy - m.v * m.v, pzz = m.zz - m.w * m.w;
            const R pxy = m.xy - m.u * m.v, pxz = m.xz - m.u * m.w, pyz = m.yz - m.v * m.w;
            const R om = relax_rate<TWO, R>(dp, nr, pxx, pyy, pzz, pxy, pxz, pyz);
#pragma unroll
            for (int q = 0; q < 10; ++q) lbc[q] = 0;
#pragma unroll
            for (int c = 0; c < 27; ++c) {
              const int p = c % 3 - 1, q = (c / 3) % 3 - 1, r = c / 9 - 1;
              if (c == 13 || !present<LAT>(p, q, r)) continue;
              if (isf_at(a.isfh, xs - p, ys - q, zsrc - r, nx, ny, nz) != 0) continue;
              const R wl = (R)weight<LAT>(p, q, r);
              const R fneq = wl * (R)4.5 *
                             (((R)(p * p) - cssq) * pxx + ((R)(q * q) - cssq) * pyy + ((R)(r * r) - cssq) * pzz +
                              (R)(2 * p * q) * pxy + (R)(2 * p * r) * pxz + (R)(2 * q * r) * pyz);
              const R fp = wl * m.P + ((R)1 - om) * fneq;
              lbc[0] += fp;
              lbc[1] += fp * (R)p; lbc[2] += fp * (R)q; lbc[3] += fp * (R)r;
              lbc[4] += fp * ((R)(p * p) - cssq); lbc[5] += fp * ((R)(q * q) - cssq); lbc[6] += fp * ((R)(r * r) - cssq);
              lbc[7] += fp * (R)(p * q); lbc[8] += fp * (R)(p * r); lbc[9] += fp * (R)(q * r);
            }
          }
        }
      }
      if (solid) {
        m.P = m.u = m.v = m.w = m.xx = m.yy = m.zz = m.xy = m.xz = m.yz = 0;
        Fx = Fy = Fz = 0;
      }
      in.P = m.P; in.u = m.u; in.v = m.v; in.w = m.w;
      in.xx = m.xx - m.u * m.u; in.yy = m.yy - m.v * m.v; in.zz = m.zz - m.w * m.w;
      in.xy = m.xy - m.u * m.v; in.xz = m.xz - m.u * m.w; in.yz = m.yz - m.v * m.w;
      in.ax = Fx; in.ay = Fy; in.az = Fz;
      in.om = relax_rate<TWO, R>(dp, nr, in.xx, in.yy, in.zz, in.xy, in.xz, in.yz);
      if constexpr (PF) {
        // every store of this iteration is issued: stream the next plane in
        lds_wait_all();
        if (zsrc < zend) {
          if (!halo_wave) dma_raw<PREC, TWO, INLINE, GI>(a, wrapg(zsrc + 1, nz), fs, off, lds0);
          if (obst) isf_next = isf_load(zsrc + 1);
        }
        if (GI && zsrc + 3 <= zend + 1) dma_phi(zsrc + 3);
        if constexpr (sizeof(T) == 2) __builtin_amdgcn_s_setprio(0);
      }
      collide<LAT, HO, R>(in, f);
      if (bb && isf == -1 && out_node && own_plane) {
        // BOUNCE_BACK (fused_LB_kernel_bb, fused:364-382): a population whose
        // upstream node is solid is replaced by this node's own opposite
        // post-collision population.  The solid node's pushed population is
        // exactly zero (its moments and force are zeroed), so the pull below
        // adds zero and the replacement is this correction.
        has_lbc = true;
        const R cssq = (R)(1.0 / 3.0);
#pragma unroll
        for (int q = 0; q < 10; ++q) lbc[q] = 0;
#pragma unroll
        for (int c = 0; c < 27; ++c) {
          const int p = c % 3 - 1, q = (c / 3) % 3 - 1, r = c / 9 - 1;
          if (c == 13 || !present<LAT>(p, q, r)) continue;
          if (isf_at(a.isfh, xs - p, ys - q, zsrc - r, nx, ny, nz) != 0) continue;
          const R fo = f[cidx(-p, -q, -r)];
          lbc[0] += fo;
          lbc[1] += fo * (R)p; lbc[2] += fo * (R)q; lbc[3] += fo * (R)r;
          lbc[4] += fo * ((R)(p * p) - cssq); lbc[5] += fo * ((R)(q * q) - cssq); lbc[6] += fo * ((R)(r * r) - cssq);
          lbc[7] += fo * (R)(p * q); lbc[8] += fo * (R)(p * r); lbc[9] += fo * (R)(q * r);
        }
      }
    }

    // ---- 2. x-halo slices through LDS -------------------------------------------
    if (halo_wave && active) {
      // left halo streams +x into lane 0, right -x into lane XW-1.  The two
      // sides are separate branches with constant register indices: a
      // select between the sides' values is turned by the compiler into a
      // lane-variable index and expanded into a 27-way compare chain.
      if (side == 0) {
#pragma unroll
        for (int r = -1; r <= 1; ++r)
#pragma unroll
          for (int q = -1; q <= 1; ++q) sHX[w][0][(q + 1) + 3 * (r + 1)] = f[cidx(1, q, r)];
      } else {
#pragma unroll
        for (int r = -1; r <= 1; ++r)
#pragma unroll
          for (int q = -1; q <= 1; ++q) sHX[w][1][(q + 1) + 3 * (r + 1)] = f[cidx(-1, q, r)];
      }
    }
    if constexpr (PF) {
      // The x-halo wave's next-plane inputs (14 scattered dwords per lane)
      // are issued here, after its collision, instead of with the row
      // waves' transfers: the issue points of a block's waves spread over
      // the iteration (fused kernel 3.91 -> 3.82 ms FP32, 3.37 -> 3.22 ms
      // FP16S at 512^3).  Its slots were read at the top of the iteration.
      if (halo_wave && zsrc < zend) {
        lds_wait_all();
        dma_raw<PREC, TWO, INLINE, GI>(a, wrapg(zsrc + 1, nz), (size_t)sx, off, lds0);
      }
    }
    __syncthreads();

    R G0[9];  // q = 0 partial sums kept in registers: [d][r]
    if (!halo_wave) {
      // ---- 3. x: shift by one lane, contract over p ---------------------------
      // the x-halo values enter as DPP fill operands, read by every lane
      // from one LDS address (a broadcast)
      const bool short_row = XW < 64;  // the right halo enters below lane 63
      R G[3][3][3];  // [d][q][r]
#pragma unroll
      for (int r = -1; r <= 1; ++r)
#pragma unroll
        for (int q = -1; q <= 1; ++q) {
          const int qr = (q + 1) + 3 * (r + 1);
          R fm_in = 0, fp_in = 0;
          if (present<LAT>(1, q, r)) fp_in = shift_from_lower(f[cidx(1, q, r)], sHX[wave][0][qr]);
          if (present<LAT>(-1, q, r)) {
            const R hr = sHX[wave][1][qr];
            fm_in = shift_from_upper(f[cidx(-1, q, r)], hr);
            if (short_row) fm_in = lane == XW - 1 ? hr : fm_in;
          }
          const R f0 = present<LAT>(0, q, r) ? f[cidx(0, q, r)] : (R)0;
          const R s = fm_in + f0 + fp_in;
          G[0][q + 1][r + 1] = s;
          G[1][q + 1][r + 1] = fp_in - fm_in;
          G[2][q + 1][r + 1] = twothird * (fp_in + fm_in) - third * f0;
        }
      // ---- 4a. y: hand the q = +-1 partial sums to the neighbouring rows ------
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          if (wave < NW - 1) sUP[wave][d * 3 + r][lane] = G[d][2][r];
          if (wave > 0) sDN[wave][d * 3 + r][lane] = G[d][0][r];
          G0[d * 3 + r] = G[d][1][r];
        }
    }
    __syncthreads();

    if (out_row) {
      // ---- 4b. y: contract over q ------------------------------------------------
      // H[d][e][r] for (d,e) in {00, 01, 02, 10, 11, 20}
      R H00[3], H01[3], H02[3], H10[3], H11[3], H20[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        {
          const R gp = sUP[wave - 1][0 * 3 + r][lane], gm = sDN[wave + 1][0 * 3 + r][lane], g0 = G0[0 * 3 + r];
          H00[r] = gm + g0 + gp;
          H01[r] = gp - gm;
          H02[r] = twothird * (gp + gm) - third * g0;
        }
        {
          const R gp = sUP[wave - 1][1 * 3 + r][lane], gm = sDN[wave + 1][1 * 3 + r][lane], g0 = G0[1 * 3 + r];
          H10[r] = gm + g0 + gp;
          H11[r] = gp - gm;
        }
        {
          const R gp = sUP[wave - 1][2 * 3 + r][lane], gm = sDN[wave + 1][2 * 3 + r][lane], g0 = G0[2 * 3 + r];
          H20[r] = gm + g0 + gp;
        }
      }
      // ---- 5. z: r = +1 -> plane zsrc+1 (hi), 0 -> zsrc (mid), -1 -> zsrc-1 (lo)
      // outputs: 0 P(000) 1 u(100) 2 v(010) 3 w(001) 4 xx(200) 5 yy(020) 6 zz(002)
      //          7 xy(110) 8 xz(101) 9 yz(011); index [0]=r-1 [1]=r0 [2]=r+1
      // plane zsrc+1 receives its first term here: assign rather than add to
      // a zero (0 + x cannot be folded for signed zeros)
      hi[0] = H00[2]; mid[0] += H00[1]; lo[0] += H00[0];
      hi[3] = H00[2]; lo[3] -= H00[0];
      hi[6] = twothird * H00[2]; mid[6] -= third * H00[1]; lo[6] += twothird * H00[0];
      hi[2] = H01[2]; mid[2] += H01[1]; lo[2] += H01[0];
      hi[9] = H01[2]; lo[9] -= H01[0];
      hi[1] = H10[2]; mid[1] += H10[1]; lo[1] += H10[0];
      hi[8] = H10[2]; lo[8] -= H10[0];
      hi[5] = H02[2]; mid[5] += H02[1]; lo[5] += H02[0];
      hi[4] = H20[2]; mid[4] += H20[1]; lo[4] += H20[0];
      hi[7] = H11[2]; mid[7] += H11[1]; lo[7] += H11[0];
      if (has_lbc) {
#pragma unroll
        for (int q = 0; q < 10; ++q) mid[q] += lbc[q];
      }
      // ---- 6. plane zsrc-1 is complete (stored here, or by the next iteration) --
      if constexpr (!PF) {
        const int zo = zsrc - 1;
        if (lane < XW && zo >= z0 && zo < zend) store_lo(zo, fs);
#pragma unroll
        for (int m = 0; m < 10; ++m) { lo[m] = mid[m]; mid[m] = hi[m]; }
      }
    }
  }
  if constexpr (PF) {
    if (out_row && lane < XW) store_lo(zend - 1, (size_t)sx);
  }
}

// One domain per launch: 1-D grid, XCD-aware tile order.
template <int LAT, bool HO, int PREC, bool TWO, bool INLINE, int TY, bool OB, bool OPEN>
__global__ void __launch_bounds__((TY + 3) * 64) fused_kernel(FusedArgs a) {
  fused_body<LAT, HO, PREC, TWO, INLINE, TY, OB, OPEN>(a, xcd_tile(blockIdx.x, fused_tiles<TY>(a)));
}

// Many domains per launch: the grid covers every domain's tiles in table
// order (domain-major); a block finds its domain by binary search over the
// tile offsets, then runs the same body on that domain's arguments.
template <int LAT, bool HO, int PREC, bool TWO, int TY>
__global__ void __launch_bounds__((TY + 3) * 64)
    fused_kernel_batch(const FusedBatchEntry* __restrict__ tab, int ndom, int total) {
  const int g = xcd_tile(blockIdx.x, total);
  int lo = 0, hi = ndom - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tab[mid].tile0 <= g) lo = mid; else hi = mid - 1;
  }
  fused_body<LAT, HO, PREC, TWO, true, TY, false>(tab[lo].a, g - tab[lo].tile0);  // batches hold no obstacles
}

// ---------------------------------------------------------------------------
template <int LAT, bool HO, int PREC, bool TWO, bool INLINE, int TY>
static hipError_t launch_t(const FusedArgs& a, hipStream_t s) {
  dim3 grid(fused_tiles<TY>(a));
  dim3 block((TY + 3) * 64);
  // open faces reach the fast path with obstacles and 32/64-bit storage only
  constexpr bool open_ok = INLINE && prec_base(PREC) <= PREC_MIXED;
  if (a.dp.obst) {
    if constexpr (open_ok) {
      if (a.sopen) {
        hipLaunchKernelGGL((fused_kernel<LAT, HO, PREC, TWO, INLINE, TY, true, true>), grid, block, 0, s, a);
        return hipGetLastError();
      }
    }
    hipLaunchKernelGGL((fused_kernel<LAT, HO, PREC, TWO, INLINE, TY, true, false>), grid, block, 0, s, a);
  } else {
    hipLaunchKernelGGL((fused_kernel<LAT, HO, PREC, TWO, INLINE, TY, false, false>), grid, block, 0, s, a);
  }
  return hipGetLastError();
}

template <int PREC, bool TWO, bool INLINE, int TY>
static hipError_t launch_lat(const FusedArgs& a, hipStream_t s) {
  if (a.dp.lattice == 19) return launch_t<19, false, PREC, TWO, INLINE, TY>(a, s);
  if (a.dp.highorder) return launch_t<27, true, PREC, TWO, INLINE, TY>(a, s);
  return launch_t<27, false, PREC, TWO, INLINE, TY>(a, s);
}

template <int PREC, bool INLINE, int TY>
static hipError_t launch_two(const FusedArgs& a, hipStream_t s) {
  if (a.dp.twocomp) return launch_lat<PREC, true, INLINE, TY>(a, s);
  return launch_lat<PREC, false, INLINE, TY>(a, s);
}

template <int PREC, bool INLINE>
static hipError_t launch_ty(const FusedArgs& a, hipStream_t s) {
  switch (a.ty) {
    case 8: return launch_two<PREC, INLINE, 8>(a, s);
    case 13: return launch_two<PREC, INLINE, 13>(a, s);
    default: return hipErrorInvalidValue;
  }
}

int fused_tile_count(const FusedArgs& a) {
  switch (a.ty) {
    case 8: return fused_tiles<8>(a);
    case 13: return fused_tiles<13>(a);
  }
  return -1;
}

template <int PREC, bool TWO, int TY>
static hipError_t batch_lat(const FusedBatchEntry* tab, int ndom, int total, int lat, bool ho, hipStream_t s) {
  dim3 grid(total), block((TY + 3) * 64);
  if (lat == 19) hipLaunchKernelGGL((fused_kernel_batch<19, false, PREC, TWO, TY>), grid, block, 0, s, tab, ndom, total);
  else if (ho) hipLaunchKernelGGL((fused_kernel_batch<27, true, PREC, TWO, TY>), grid, block, 0, s, tab, ndom, total);
  else hipLaunchKernelGGL((fused_kernel_batch<27, false, PREC, TWO, TY>), grid, block, 0, s, tab, ndom, total);
  return hipGetLastError();
}

template <int PREC>
static hipError_t batch_prec(const FusedBatchEntry* tab, int ndom, int total, int lat, bool ho, bool two, int ty,
                             hipStream_t s) {
  if (ty == 8) return two ? batch_lat<PREC, true, 8>(tab, ndom, total, lat, ho, s) : batch_lat<PREC, false, 8>(tab, ndom, total, lat, ho, s);
  if (ty == 13) return two ? batch_lat<PREC, true, 13>(tab, ndom, total, lat, ho, s) : batch_lat<PREC, false, 13>(tab, ndom, total, lat, ho, s);
  return hipErrorInvalidValue;
}

hipError_t launch_fused_batch(const FusedBatchEntry* dev_tab, int ndom, int total_tiles, int prec, int lattice,
                              bool highorder, bool two, int ty, hipStream_t s) {
  switch (prec) {
    case PREC_FP16S | PREC_DEFSCALE:
      return batch_prec<PREC_FP16S | PREC_DEFSCALE>(dev_tab, ndom, total_tiles, lattice, highorder, two, ty, s);
    case PREC_BF16S | PREC_DEFSCALE:
      return batch_prec<PREC_BF16S | PREC_DEFSCALE>(dev_tab, ndom, total_tiles, lattice, highorder, two, ty, s);
    case PREC_FP64: return batch_prec<PREC_FP64>(dev_tab, ndom, total_tiles, lattice, highorder, two, ty, s);
    case PREC_FP32: return batch_prec<PREC_FP32>(dev_tab, ndom, total_tiles, lattice, highorder, two, ty, s);
    case PREC_MIXED: return batch_prec<PREC_MIXED>(dev_tab, ndom, total_tiles, lattice, highorder, two, ty, s);
    case PREC_FP16S: return batch_prec<PREC_FP16S>(dev_tab, ndom, total_tiles, lattice, highorder, two, ty, s);
    case PREC_BF16S: return batch_prec<PREC_BF16S>(dev_tab, ndom, total_tiles, lattice, highorder, two, ty, s);
  }
  return hipErrorInvalidValue;
}

bool default_scales(const DevParams& dp) {
  for (int i = 0; i < 10; ++i)
    if (dp.scale[i] != def_scale(i)) return false;
  return true;
}

hipError_t launch_fused(const FusedArgs& a, int prec, bool inline_moments, hipStream_t s) {
  if (inline_moments) {
    const bool ds = default_scales(a.dp);
    switch (prec) {
      case PREC_FP64: return launch_ty<PREC_FP64, true>(a, s);
      case PREC_FP32: return launch_ty<PREC_FP32, true>(a, s);
      case PREC_MIXED: return launch_ty<PREC_MIXED, true>(a, s);
      case PREC_FP16S:
        return ds ? launch_ty<PREC_FP16S | PREC_DEFSCALE, true>(a, s) : launch_ty<PREC_FP16S, true>(a, s);
      case PREC_BF16S:
        return ds ? launch_ty<PREC_BF16S | PREC_DEFSCALE, true>(a, s) : launch_ty<PREC_BF16S, true>(a, s);
    }
  } else {
    switch (prec) {
      case PREC_FP64: return launch_ty<PREC_FP64, false>(a, s);
      case PREC_FP32: return launch_ty<PREC_FP32, false>(a, s);
      case PREC_MIXED: return launch_ty<PREC_MIXED, false>(a, s);
    }
  }
  return hipErrorInvalidValue;
}

}  // namespace lbf
