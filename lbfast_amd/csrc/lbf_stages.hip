// lbf_stages.hip -- the reference's per-stage kernels for gfx950 (one thread
// per node, row-interleaved field sets, in-kernel periodic wrap), the tiled phase-field
// prep kernel of the fast path, and the encode/decode/reduction helpers.
//
//   k_norm        compute_norm_interface_kernel (lb_cuda_auxfields_module.f90:19-139)
//   k_div         compute_div_theta_n_kernel    (lb_cuda_auxfields_module.f90:1085-1164)
//   k_phi_bnd     PHI_int_boundary_kernel       (lb_cuda_boundary_module.f90:218-459)
//   k_moments     moments_LB_kernel             (lb_cuda_moments_module.f90:19-341)
//   k_update_phi  update_phifields_kernel       (lb_cuda_update_phi_module.f90:19-104)
//   k_lb_int_bnd  LB_int_boundary_kernel        (lb_cuda_boundary_module.f90:19-216)
//   k_phase_prep  compute_norm + compute_div_theta_n of the fast path in one
//                 z-marching pass: phi planes and theta planes live in LDS,
//                 the output per fluid node is grad phi (as n*|g|), lap phi
//                 and div(theta n) -- what the fused kernel's inline moments
//                 and phase update consume.
// The fused stage is the z-marching kernel of lbf_fused.hip.
#include <mutex>

#include "lbf_common.h"
#include "lbf_kernels.h"
#include "lbf_physics.h"

namespace lbf {

static constexpr int BS = 256;

// reference velocity order (vars_module.f90:319-338)
__constant__ int cEX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, -1, 1, 1, -1};
__constant__ int cEY[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, -1, 1, -1, 1};
__constant__ int cEZ[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1};

__device__ __forceinline__ bool node_of(size_t t, const DevParams& dp, int& x, int& y, int& z) {
  if (t >= dp.nn) return false;
  x = (int)(t % dp.nx);
  const size_t r = t / dp.nx;
  y = (int)(r % dp.ny);
  z = (int)(r / dp.ny);
  return true;
}

#define NB(f, a, b, c) ((R)(f)[nidx(wrapg(x + (a), nx), wrapg(y + (b), ny), wrapg(z + (c), nz), nx, ny)])
// neighbour in a field set of F fields; f points at the wanted field's offset
#define NBS(f, F, a, b, c) ((R)(f)[sbase((F), wrapg(x + (a), nx), wrapg(y + (b), ny), wrapg(z + (c), nz), sx, ny)])

// isotropic D3Q27 gradient / laplacian of a field at (0,0,0), given an
// accessor F(a,b,c); expression order of lb_cuda_auxfields_module.f90:72-130.
// DX/DY/DZ are the paired-difference sums without the leading factor 3.
#define DX(F)                                                                                            \
  (p1 * (F(1, 0, 0) - F(-1, 0, 0)) +                                                                     \
   p2 * ((F(1, 1, 0) - F(-1, -1, 0)) + (F(1, -1, 0) - F(-1, 1, 0)) + (F(1, 0, 1) - F(-1, 0, -1)) +        \
         (F(1, 0, -1) - F(-1, 0, 1))) +                                                                  \
   p3 * ((F(1, 1, 1) - F(-1, -1, -1)) + (F(1, -1, -1) - F(-1, 1, 1)) + (F(1, -1, 1) - F(-1, 1, -1)) +     \
         (F(1, 1, -1) - F(-1, -1, 1))))
#define DY(F)                                                                                            \
  (p1 * (F(0, 1, 0) - F(0, -1, 0)) +                                                                     \
   p2 * ((F(1, 1, 0) - F(-1, -1, 0)) + (F(-1, 1, 0) - F(1, -1, 0)) + (F(0, 1, 1) - F(0, -1, -1)) +        \
         (F(0, 1, -1) - F(0, -1, 1))) +                                                                  \
   p3 * ((F(1, 1, 1) - F(-1, -1, -1)) + (F(-1, 1, -1) - F(1, -1, 1)) + (F(1, 1, -1) - F(-1, -1, 1)) +     \
         (F(-1, 1, 1) - F(1, -1, -1))))
#define DZ(F)                                                                                            \
  (p1 * (F(0, 0, 1) - F(0, 0, -1)) +                                                                     \
   p2 * ((F(1, 0, 1) - F(-1, 0, -1)) + (F(-1, 0, 1) - F(1, 0, -1)) + (F(0, 1, 1) - F(0, -1, -1)) +        \
         (F(0, -1, 1) - F(0, 1, -1))) +                                                                  \
   p3 * ((F(1, 1, 1) - F(-1, -1, -1)) + (F(-1, -1, 1) - F(1, 1, -1)) + (F(1, -1, 1) - F(-1, 1, -1)) +     \
         (F(-1, 1, 1) - F(1, -1, -1))))
#define GRAD_X(F) (THREE * DX(F))
#define GRAD_Y(F) (THREE * DY(F))
#define GRAD_Z(F) (THREE * DZ(F))
#define LAPL(F, c)                                                                                                  \
  ((TWO * THREE) *                                                                                                  \
   ((c) * (p0 - ONE) +                                                                                              \
    (p1 * (F(1, 0, 0) + F(-1, 0, 0) + F(0, 1, 0) + F(0, -1, 0) + F(0, 0, 1) + F(0, 0, -1)) +                        \
     p2 * ((F(1, 1, 0) + F(-1, -1, 0)) + (F(1, -1, 0) + F(-1, 1, 0)) + (F(1, 0, 1) + F(-1, 0, -1)) +                \
           (F(1, 0, -1) + F(-1, 0, 1)) + (F(0, 1, 1) + F(0, -1, -1)) + (F(0, 1, -1) + F(0, -1, 1))) +              \
     p3 * ((F(1, 1, 1) + F(-1, -1, -1)) + (F(1, -1, -1) + F(-1, 1, 1)) + (F(1, -1, 1) + F(-1, 1, -1)) +            \
           (F(1, 1, -1) + F(-1, -1, 1))))))

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
  const R gx = GRAD_X(F), gy = GRAD_Y(F), gz = GRAD_Z(F);
  const R mg = sqrt(gx * gx + gy * gy + gz * gz);
  const R eps = (R)1.0e-9;
  const R c = F(0, 0, 0);
  const R lap = LAPL(F, c);
#undef F
  const size_t n = nidx(x, y, z, nx, ny);
  const int sx = dp.sx;
  const size_t n7 = sbase(7, x, y, z, sx, ny);
  TA* aux = (TA*)a.aux;
  aux[n7] = (TA)(gx / (mg + eps));
  aux[n7 + sx] = (TA)(gy / (mg + eps));
  aux[n7 + 2 * sx] = (TA)(gz / (mg + eps));
  aux[n7 + 3 * sx] = (TA)mg;
  aux[n7 + 4 * sx] = (TA)(c * (ONE - c) * (gx / (mg + eps)));
  aux[n7 + 5 * sx] = (TA)(c * (ONE - c) * (gy / (mg + eps)));
  aux[n7 + 6 * sx] = (TA)(c * (ONE - c) * (gz / (mg + eps)));
  ((TA*)a.lap)[n] = (TA)lap;
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
  const int sx = dp.sx;
  const TA* ax = (const TA*)a.aux + 4 * sx;
  const TA* ay = (const TA*)a.aux + 5 * sx;
  const TA* az = (const TA*)a.aux + 6 * sx;
  const R p1 = (R)(2.0 / 27.0), p2 = (R)(1.0 / 54.0), p3 = (R)(1.0 / 216.0);
#define X(a_, b_, c_) NBS(ax, 7, a_, b_, c_)
#define Y(a_, b_, c_) NBS(ay, 7, a_, b_, c_)
#define Z(a_, b_, c_) NBS(az, 7, a_, b_, c_)
  const R d = (DX(X) + DY(Y) + DZ(Z)) * (R)3;
#undef X
#undef Y
#undef Z
  ((TA*)a.dvt)[nidx(x, y, z, nx, ny)] = (TA)d;
}

// ---------------------------------------------------------------------------
// Fast-path phase-field prep: z-marching over a 64 x PTY tile.  Iteration k
// loads phi(k+1), computes theta(k) on tile+1 (and grad/lap for the tile),
// then div theta at plane k-1 from the theta ring.  Outputs (fluid nodes):
// pre = {n_x |g|, n_y |g|, n_z |g|, lap phi, div(phi(1-phi) n)}.  theta is 0 at
// non-fluid nodes: the reference never writes aux there and reads the
// zero-initialised allocation.
constexpr int PTX = 64, PTY = 8, PNT = PTX * PTY;

template <int PREC>
__global__ void __launch_bounds__(PNT) k_phase_prep(StageArgs a, int tz) {
  typedef typename Codec<PREC>::R R;
  typedef typename Aux<PREC>::T TA;
  constexpr int SX = PTX + 4, SY = PTY + 4;   // phi plane with a 2-node halo
  constexpr int TX = PTX + 2, TYY = PTY + 2;  // theta plane with a 1-node halo
  constexpr int NRING = TX * TYY - PTX * PTY; // theta halo-ring nodes per plane
  static_assert(NRING <= PNT, "one ring node per thread at most");
  __shared__ R sphi[3][SY][SX];
  __shared__ R sth[3][3][TYY][TX];

  const DevParams& dp = a.dp;
  const int nx = dp.nx, ny = dp.ny, nz = dp.nz;
  const size_t nxny = (size_t)nx * ny;
  const int sx = dp.sx;
  const int tid = threadIdx.x;
  const int ntx = (nx + PTX - 1) / PTX, nty = (ny + PTY - 1) / PTY, ntz = (nz + tz - 1) / tz;
  const int tile = xcd_tile(blockIdx.x, ntx * nty * ntz);
  const int x0 = (tile % ntx) * PTX, y0 = ((tile / ntx) % nty) * PTY;
  const int z0 = (tile / (ntx * nty)) * tz, zend = min(z0 + tz, nz);
  const TA* phi = (const TA*)a.phiS;
  const R p1 = (R)(2.0 / 27.0), p2 = (R)(1.0 / 54.0), p3 = (R)(1.0 / 216.0), p0 = (R)(8.0 / 27.0);
  const R THREE = 3, ONE = 1, TWO = 2;
  const R eps = (R)1.0e-9;

  auto slot3 = [](int k) { return ((k % 3) + 3) % 3; };
  auto load_plane = [&](int z) {
    const int zw = ((z % nz) + nz) % nz;
    const int sl = slot3(z);
    const TA* pz = phi + (size_t)zw * nxny;
    for (int t = tid; t < SX * SY; t += PNT) {
      const int ix = t % SX, iy = t / SX;
      const int gx = (((x0 - 2 + ix) % nx) + nx) % nx, gy = (((y0 - 2 + iy) % ny) + ny) % ny;
      sphi[sl][iy][ix] = (R)pz[(size_t)gx + (size_t)nx * gy];
    }
  };
  // theta (and, for an interior node, grad/lap) of theta-plane node (jx,jy)
  // at plane k from the phi slots; returns true if the node is fluid
  auto theta_node = [&](int jx, int jy, int k, bool interior, R& gfx, R& gfy, R& gfz, R& lap) {
    const int sm = slot3(k - 1), s0 = slot3(k), sp = slot3(k + 1);
    const int8_t isf = isf_at(a.isfh, x0 - 1 + jx, y0 - 1 + jy, k, nx, ny, nz);
    R thx = 0, thy = 0, thz = 0;
    if (isf == 1 || isf == -1) {
#define F(a_, b_, c_) sphi[(c_) < 0 ? sm : ((c_) > 0 ? sp : s0)][jy + 1 + (b_)][jx + 1 + (a_)]
      const R gx = GRAD_X(F), gy = GRAD_Y(F), gz = GRAD_Z(F);
      const R mg = sqrt(gx * gx + gy * gy + gz * gz);
      const R im = rcp(mg + eps);
      const R c = F(0, 0, 0);
      const R nxv = gx * im, nyv = gy * im, nzv = gz * im;
      thx = c * (ONE - c) * nxv; thy = c * (ONE - c) * nyv; thz = c * (ONE - c) * nzv;
      if (interior) {
        // the reference re-reads n and |g| from storage precision
        const R mgs = (R)(TA)mg;
        gfx = (R)(TA)nxv * mgs; gfy = (R)(TA)nyv * mgs; gfz = (R)(TA)nzv * mgs;
        lap = LAPL(F, c);
      }
#undef F
    }
    sth[s0][0][jy][jx] = (R)(TA)thx;
    sth[s0][1][jy][jx] = (R)(TA)thy;
    sth[s0][2][jy][jx] = (R)(TA)thz;
  };

  load_plane(z0 - 2);
  load_plane(z0 - 1);

  // my interior node (tx,ty) is theta node (tx+1, ty+1); ring node `tid` if tid < NRING
  const int tx = tid % PTX, ty = tid / PTX;
  const int xg = x0 + tx, yg = y0 + ty;
  const bool mine = xg < nx && yg < ny;
  int rjx = 0, rjy = 0;
  if (tid < NRING) {
    if (tid < TX) { rjx = tid; rjy = 0; }
    else if (tid < 2 * TX) { rjx = tid - TX; rjy = TYY - 1; }
    else if (tid < 2 * TX + PTY) { rjx = 0; rjy = tid - 2 * TX + 1; }
    else { rjx = TX - 1; rjy = tid - 2 * TX - PTY + 1; }
  }
  R gfx_c = 0, gfy_c = 0, gfz_c = 0, lap_c = 0;  // my node at plane k-1
  for (int k = z0 - 1; k <= zend; ++k) {
    load_plane(k + 1);
    __syncthreads();
    R gfx_n = 0, gfy_n = 0, gfz_n = 0, lap_n = 0;
    theta_node(tx + 1, ty + 1, k, true, gfx_n, gfy_n, gfz_n, lap_n);
    if (tid < NRING) {
      R d0, d1, d2, d3;
      theta_node(rjx, rjy, k, false, d0, d1, d2, d3);
    }
    __syncthreads();
    // div theta at plane k-1 for my node
    const int zo = k - 1;
    if (mine && zo >= z0 && zo < zend) {
      const int8_t isf = isf_at(a.isfh, xg, yg, zo, nx, ny, nz);
      if (isf == 1 || isf == -1) {
        const int tm = slot3(zo - 1), t0 = slot3(zo), tp = slot3(zo + 1);
        const int jx = tx + 1, jy = ty + 1;
#define X(a_, b_, c_) sth[(c_) < 0 ? tm : ((c_) > 0 ? tp : t0)][0][jy + (b_)][jx + (a_)]
#define Y(a_, b_, c_) sth[(c_) < 0 ? tm : ((c_) > 0 ? tp : t0)][1][jy + (b_)][jx + (a_)]
#define Z(a_, b_, c_) sth[(c_) < 0 ? tm : ((c_) > 0 ? tp : t0)][2][jy + (b_)][jx + (a_)]
        const R d = (DX(X) + DY(Y) + DZ(Z)) * THREE;
#undef X
#undef Y
#undef Z
        const size_t o = sbase(5, xg, yg, zo, sx, ny);
        TA* pre = (TA*)a.pre;
        pre[o] = (TA)gfx_c;
        pre[o + sx] = (TA)gfy_c;
        pre[o + 2 * sx] = (TA)gfz_c;
        pre[o + 3 * sx] = (TA)lap_c;
        pre[o + 4 * sx] = (TA)d;
      }
    }
    gfx_c = gfx_n; gfy_c = gfy_n; gfz_c = gfz_n; lap_c = lap_n;
  }
}

// ---------------------------------------------------------------------------
// Streaming fast-path phase-field prep (float phase-field storage).
//
// The stencils of k_norm / k_div (lb_cuda_auxfields_module.f90:72-130,
// 1110-1160) in their separable form (part3 / grad_lap in lbf_common.h):
// each plane contributes three 9-point partial sums per node, and every
// thread keeps the partials of the last planes of its own nodes in
// registers as it marches z.  That reads 9 phi values per node from LDS (27
// direct) and 21 theta values (54 direct).  The sums run in a different
// order than the direct stencils, so results agree with them to rounding,
// not bit for bit.
//
// A block owns a 64-column tile (outputs x0 .. x0+63, aligned with the field
// rows, so every store is a whole-row segment) by QY rows and marches z.
// Its QW row waves hold two theta rows each; one more wave computes theta on
// the two edge columns x0-1 and x0+64 that the tile's divergence needs.
// phi rows are kept in LDS with a 2-column halo each side (column c holds
// x0-2+c, wrapped): the row's 64 lanes and two 2-lane transfers for the
// halo columns stream in by LDS DMA two planes ahead, into a ring of four
// planes; theta lives in a ring of two planes of 66 columns (x0-1 ..
// x0+64).  Iteration k computes div theta of plane k-2, then reads phi
// plane k+1 and computes grad/lap/theta of plane k, with one barrier per
// plane.  isfluid is read only with internal obstacles.  With a.div_only the gradient and Laplacian are not stored
// (the fused kernel forms them itself from phi); only div theta is.
constexpr int QY = 12, QW = 7;
constexpr int QR = QY + 4;                 // phi rows of a tile (2-row halo)
constexpr int QT = QY + 2;                 // theta rows (1-row halo)
constexpr int QC = 64 + 4;                 // phi columns x0-2 .. x0+65
constexpr int TC = 64 + 2;                 // theta columns x0-1 .. x0+64
constexpr int QDMA = (QR + QW - 1) / QW;   // phi rows each row wave streams per plane
constexpr int QNT = (QW + 1) * 64;         // threads: row waves + the edge-column wave

__host__ __device__ inline int prep_tiles_z(const DevParams& dp, int tz) {
  return ((dp.nx + 63) / 64) * ((dp.ny + QY - 1) / QY) * ((dp.nz + tz - 1) / tz);
}

// DO: div theta only (a.div_only); the gradient outputs and the registers
// that hold them are compiled out
template <int PREC, bool DO>
__device__ __forceinline__ void prep_body_z(const StageArgs& a, int tz, int tile) {
  typedef typename Codec<PREC>::R R;
  typedef typename Aux<PREC>::T TA;
  static_assert(sizeof(TA) == 4, "float phase-field storage");
  __shared__ float sphi[4][QR + 1][QC];  // row QR absorbs the padding transfers
  __shared__ R sth[2][3][QT][TC];

  const DevParams& dp = a.dp;
  const int nx = dp.nx, ny = dp.ny, nz = dp.nz;
  const size_t nxny = (size_t)nx * ny;
  const int sx = dp.sx;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const bool edge_wave = wave == QW;
  const int ntx = (nx + 63) / 64, nty = (ny + QY - 1) / QY;
  const int x0 = (tile % ntx) * 64, y0 = ((tile / ntx) % nty) * QY;
  const int z0 = (tile / (ntx * nty)) * tz, zend = min(z0 + tz, nz);
  const int XW = min(64, nx - x0);
  const uint32_t pb = (uint32_t)(nxny * sizeof(TA));
  const TA* phi = (const TA*)a.phiS;
  TA* pre = (TA*)a.pre;
  const R FOUR = 4, ONE = 1, HALF = 0.5, G3 = (R)(3.0 / 216.0);
  const R eps = (R)1.0e-9;
  constexpr bool store_g = !DO;
  auto wrapx = [&](int x) { return ((x % nx) + nx) % nx; };

  // theta column of this lane: row waves j = lane + 1 (x = x0 + lane); the
  // edge wave's lane 2t + s computes theta row t at j = 0 (s = 0) or 65
  int tj = lane + 1, trow0 = 0;
  bool edge_active = true;
  if (edge_wave) {
    edge_active = lane < 2 * QT;
    tj = (lane & 1) ? TC - 1 : 0;
    trow0 = edge_active ? (lane >> 1) : 0;
  }
  const int xt = x0 - 1 + tj;            // x of this lane's theta column (unwrapped)
  const int xo = x0 + lane;              // output column of a row-wave lane
  const bool out_lane = !edge_wave && lane < XW;

  // DMA: row wave w streams phi rows w + QW*i: its 64 lanes to columns
  // 2..65, lanes 0-1 to columns 0-1 (x0-2, x0-1) and to 66-67 (x0+64,
  // x0+65); rows past the tile go to the padding row
  // row byte offsets are wave-uniform (scalar); the lane's column offsets
  // are added when a transfer is issued
  uint32_t rowbyte[QDMA], rowlds[QDMA];
  const uint32_t xm = (uint32_t)wrapx(x0 + lane) * 4u;
  const uint32_t xlh = (uint32_t)wrapx(x0 - 2 + (lane & 1)) * 4u, xrh = (uint32_t)wrapx(x0 + 64 + (lane & 1)) * 4u;
#pragma unroll
  for (int i = 0; i < QDMA; ++i) {
    const int r = wave + QW * i;
    const int yw = (((y0 - 2 + (r < QR ? r : 0)) % ny) + ny) % ny;
    rowbyte[i] = __builtin_amdgcn_readfirstlane((uint32_t)yw * (uint32_t)nx * (uint32_t)sizeof(TA));
    rowlds[i] = __builtin_amdgcn_readfirstlane((uint32_t)(r < QR ? r : QR) * (uint32_t)QC * 4u);
  }
  const uint32_t lds_phi = lds_addr(&sphi[0][0][0]);
  auto dma_plane = [&](int z) {
    if (edge_wave) return;
    const int zw = ((z % nz) + nz) % nz;
    const Rsrc rs = mkrsrc(phi + (size_t)zw * nxny, pb);
    const uint32_t base = lds_phi + (uint32_t)(z & 3) * (QR + 1) * QC * 4u;
#pragma unroll
    for (int i = 0; i < QDMA; ++i) dma_dword<8u>(rs, rowbyte[i] + xm, base + rowlds[i]);
    // the halo columns: two lanes each (the transfers of the other lanes are
    // masked off); three transfers per row in all
    if (lane < 2) {
#pragma unroll
      for (int i = 0; i < QDMA; ++i) {
        dma_dword<0u>(rs, rowbyte[i] + xlh, base + rowlds[i]);
        dma_dword<66u * 4u>(rs, rowbyte[i] + xrh, base + rowlds[i]);
      }
    }
  };
  auto fluid = [&](int x, int y, int z) -> bool {
    if (!dp.obst) return true;
    const int8_t v = isf_at(a.isfh, x, y, z, nx, ny, nz);
    return v == 1 || v == -1;
  };

  // per theta row h of this lane: phi partials of planes k-1 and k (pm, p0;
  // p0.c is phi at the node in plane k); per div row: theta partials of
  // planes k-3 and k-2 (em, e0)
  Part3<R> pm[2], p0[2], em[2], e0[2];
  // outputs held for one iteration (see below)
  R hg[2][4], hd[2];
  size_t og[2] = {0, 0}, od[2] = {0, 0};
  bool st_g[2] = {false, false}, st_d[2] = {false, false};

  dma_plane(z0 - 2);
  dma_plane(z0 - 1);
  // Iteration k: div theta of plane k-2 from the theta of planes k-3 .. k-1,
  // then phi partials of plane k+1 and theta of plane k.  One barrier per
  // plane: theta(k-1), written before it, is read after it; theta(k) goes to
  // the other slot of the two-plane ring.  Iterations z0-3 and z0-2 only fill
  // the partials of planes z0-2, z0-1; iteration zend+1 only finishes the div.
  for (int k = z0 - 3; k <= zend + 1; ++k) {
    // phi(k+1) has landed once only the newest plane's 3*QDMA transfers are outstanding
    if (k + 2 <= zend + 1)
      asm volatile("s_waitcnt vmcnt(%0)" ::"n"(3 * QDMA) : "memory");
    else
      dma_wait_all();
    __syncthreads();

    // ---- the previous iteration's outputs ---------------------------------------
    // Stores are issued here, a full plane before the next wait on the
    // load/store counter, instead of just ahead of it.
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (st_g[h]) {
        pre[og[h]] = (TA)hg[h][0];
        pre[og[h] + sx] = (TA)hg[h][1];
        pre[og[h] + 2 * sx] = (TA)hg[h][2];
        pre[og[h] + 3 * sx] = (TA)hg[h][3];
      }
      if (st_d[h]) pre[od[h] + 4 * sx] = (TA)hd[h];
      st_g[h] = false;
      st_d[h] = false;
    }

    // ---- theta partials of plane k-1; div theta of plane k-2 on theta rows
    // 1 + wave, 1 + wave + QW -----------------------------------------------------------
    const int zo = k - 2;
    if (!edge_wave && k - 1 >= z0 - 1 && k - 1 <= zend) {
      const int tk = (k - 1) & 1;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int t = 1 + wave + QW * h;
        if (t > QY) continue;
        R w[3][3];
        Part3<R> ep;
#pragma unroll
        for (int bb = 0; bb < 3; ++bb)
#pragma unroll
          for (int aa = 0; aa < 3; ++aa) w[bb][aa] = sth[tk][0][t - 1 + bb][tj - 1 + aa];
        ep.dx = (w[0][2] - w[0][0]) + FOUR * (w[1][2] - w[1][0]) + (w[2][2] - w[2][0]);
#pragma unroll
        for (int bb = 0; bb < 3; ++bb)
#pragma unroll
          for (int aa = 0; aa < 3; ++aa) w[bb][aa] = sth[tk][1][t - 1 + bb][tj - 1 + aa];
        ep.dy = (w[2][0] - w[0][0]) + FOUR * (w[2][1] - w[0][1]) + (w[2][2] - w[0][2]);
#pragma unroll
        for (int bb = 0; bb < 3; ++bb)
#pragma unroll
          for (int aa = 0; aa < 3; ++aa) w[bb][aa] = sth[tk][2][t - 1 + bb][tj - 1 + aa];
        {
          // smoothing of (theta_z - centre), as part3
          const R ct = w[1][1];
          ep.s = ((w[0][0] - ct) + FOUR * (w[0][1] - ct) + (w[0][2] - ct)) +
                 FOUR * ((w[1][0] - ct) + (w[1][2] - ct)) + ((w[2][0] - ct) + FOUR * (w[2][1] - ct) + (w[2][2] - ct));
          ep.c = ct;
        }
        if (zo >= z0 && zo < zend) {
          const R d = G3 * ((em[h].dx + FOUR * e0[h].dx + ep.dx) + (em[h].dy + FOUR * e0[h].dy + ep.dy) +
                            (ep.s - em[h].s)) + HALF * (ep.c - em[h].c);
          const int y = y0 - 1 + t;
          if (out_lane && y < ny && fluid(xo, y, zo)) {
            hd[h] = d;
            od[h] = sbase(5, xo, y, zo, sx, ny);
            st_d[h] = true;
          }
        }
        em[h] = e0[h];
        e0[h] = ep;
      }
    }

    // ---- phi partials of plane k+1; grad / lap / theta of plane k -----------------
    // row waves: theta rows wave, wave + QW (columns 1..64); edge wave: rows
    // lane >> 1 at columns 0 and 65
    static_assert(QT == 2 * QW, "two theta rows per wave");
    if (k + 1 <= zend + 1) {
      const int sp = (k + 1) & 3, tk = k & 1;
      const bool have_theta = k >= z0 - 1;
      const bool own = k >= z0 && k < zend;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (edge_wave && h == 1) break;  // wave-uniform: the edge wave has one row per lane
        const int t = edge_wave ? trow0 : wave + QW * h;
        R v[3][3];
#pragma unroll
        for (int bb = 0; bb < 3; ++bb)
#pragma unroll
          for (int aa = 0; aa < 3; ++aa) v[bb][aa] = (R)sphi[sp][t + bb][tj + aa];
        const Part3<R> pp = part3(v);
        if (have_theta) {
          const R c = p0[h].c;
          R gx, gy, gz, lap;
          grad_lap(pm[h], p0[h], pp, gx, gy, gz, lap);
          const int y = y0 - 1 + t;
          R thx = 0, thy = 0, thz = 0;
          if (fluid(xt, y, k)) {
            const R mg = sqrt(gx * gx + gy * gy + gz * gz);
            const R im = rcp(mg + eps);
            const R nxv = gx * im, nyv = gy * im, nzv = gz * im;
            thx = c * (ONE - c) * nxv; thy = c * (ONE - c) * nyv; thz = c * (ONE - c) * nzv;
            if (store_g && own && out_lane && t >= 1 && t <= QY && y < ny) {
              // the reference re-reads n and |g| from storage precision
              const R mgs = (R)(TA)mg;
              hg[h][0] = (R)(TA)nxv * mgs;
              hg[h][1] = (R)(TA)nyv * mgs;
              hg[h][2] = (R)(TA)nzv * mgs;
              hg[h][3] = lap;
              og[h] = sbase(5, xo, y, k, sx, ny);
              st_g[h] = true;
            }
          }
          if (edge_active) {
            sth[tk][0][t][tj] = (R)(TA)thx;
            sth[tk][1][t][tj] = (R)(TA)thy;
            sth[tk][2][t][tj] = (R)(TA)thz;
          }
        }
        pm[h] = p0[h];
        p0[h] = pp;
      }
    }
    // slot (k+3)&3 held plane k-1, last read in iteration k-2
    if (k + 3 <= zend + 1) dma_plane(k + 3);
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (st_g[h]) {
      pre[og[h]] = (TA)hg[h][0];
      pre[og[h] + sx] = (TA)hg[h][1];
      pre[og[h] + 2 * sx] = (TA)hg[h][2];
      pre[og[h] + 3 * sx] = (TA)hg[h][3];
    }
    if (st_d[h]) pre[od[h] + 4 * sx] = (TA)hd[h];
  }
}

template <int PREC, bool DO>
__global__ void __launch_bounds__(QNT) k_phase_prep_z(StageArgs a, int tz) {
  prep_body_z<PREC, DO>(a, tz, xcd_tile(blockIdx.x, prep_tiles_z(a.dp, tz)));
}

// many domains per launch (see fused_kernel_batch)
template <int PREC, bool DO>
__global__ void __launch_bounds__(QNT)
    k_phase_prep_z_batch(const PrepBatchEntry* __restrict__ tab, int ndom, int total) {
  const int g = xcd_tile(blockIdx.x, total);
  int lo = 0, hi = ndom - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tab[mid].tile0 <= g) lo = mid; else hi = mid - 1;
  }
  prep_body_z<PREC, DO>(tab[lo].a, tab[lo].tz, g - tab[lo].tile0);
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
  const int sx = dp.sx;
  TA* ph = (TA*)a.phiS;
  const size_t n = nidx(x, y, z, nx, ny);
  R ghost = (R)ph[n];
  R wnx = 0, wny = 0, wnz = 0;
  for (int l = 1; l < nl; ++l) {
    const int d2 = cEX[l] * cEX[l] + cEY[l] * cEY[l] + cEZ[l] * cEZ[l];
    if (isf_at(a.isfh, x + cEX[l], y + cEY[l], z + cEZ[l], nx, ny, nz) == 0) continue;
    const R s = d2 == 1 ? (R)1 : (d2 == 2 ? half_val : third_val);
    wnx = wnx + s * (-(R)cEX[l]);
    wny = wny + s * (-(R)cEY[l]);
    wnz = wnz + s * (-(R)cEZ[l]);
  }
  const R wn = sqrt(wnx * wnx + wny * wny + wnz * wnz);
  if (wn < eps_norm) return;
  wnx = wnx / wn; wny = wny / wn; wnz = wnz / wn;
  R best = (R)-1.0e30;
  int dx = 0, dy = 0, dz = 0;
  bool have = false;
  R bsx = 0, bsy = 0, bsz = 0;
  for (int l = 1; l < nl; ++l) {
    const int d2 = cEX[l] * cEX[l] + cEY[l] * cEY[l] + cEZ[l] * cEZ[l];
    const int xi = x + cEX[l], yj = y + cEY[l], zk = z + cEZ[l];
    const int8_t v = isf_at(a.isfh, xi, yj, zk, nx, ny, nz);
    if (v != 1 && v != -1) continue;
    const R s = d2 == 1 ? (R)1 : (d2 == 2 ? inv_sqrt2 : inv_sqrt3);
    const R sx = -(R)cEX[l], sy = -(R)cEY[l], sz = -(R)cEZ[l];
    const R score = (sx * wnx + sy * wny + sz * wnz) * s;
    if (score > best) {
      best = score; dx = xi; dy = yj; dz = zk; bsx = sx; bsy = sy; bsz = sz; have = true;
    }
  }
  // the reference skips a donor with a zero (1-based) coordinate, i.e. one in
  // the lower halo (boundary:395); 0-based that is -1
  if (!have || dx == -1 || dy == -1 || dz == -1) return;
  // a donor past an upper non-periodic face (isfluid 1 there) lies in the
  // reference's halo, which nothing writes on such an axis: phi and its
  // gradient read zero, so the ghost is 0 + 0 (the donor search picks such a
  // node only where two walls meet: on a plane wall the halo's and the
  // fluid's pulls on the wall normal cancel)
  if ((!dp.pbc[0] && dx == nx) || (!dp.pbc[1] && dy == ny) || (!dp.pbc[2] && dz == nz)) {
    ph[n] = (TA)0;
    return;
  }
  const int dxw = wrapg(dx, nx), dyw = wrapg(dy, ny), dzw = wrapg(dz, nz);
  const size_t dn = nidx(dxw, dyw, dzw, nx, ny);
  const R pf = (R)ph[dn];
  R gx, gy, gz;
  if (a.gf_from_pre) {
    const TA* pre = (const TA*)a.pre;
    const size_t d5 = sbase(5, dxw, dyw, dzw, sx, ny);
    gx = (R)pre[d5]; gy = (R)pre[d5 + sx]; gz = (R)pre[d5 + 2 * sx];
  } else {
    const TA* aux = (const TA*)a.aux;
    const size_t d7 = sbase(7, dxw, dyw, dzw, sx, ny);
    const R mg = (R)aux[d7 + 3 * sx];
    gx = (R)aux[d7] * mg; gy = (R)aux[d7 + sx] * mg; gz = (R)aux[d7 + 2 * sx] * mg;
  }
  const R gdw = gx * wnx + gy * wny + gz * wnz;
  const R tx = gx - gdw * wnx, ty = gy - gdw * wny, tz = gz - gdw * wnz;
  const R tm = sqrt(tx * tx + ty * ty + tz * tz);
  const R th = ((R)180 - dp.ph<R>().wettab) * pi / (R)180;
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
  const int sx = dp.sx;
  const size_t n10 = sbase(10, x, y, z, sx, ny), n3 = sbase(3, x, y, z, sx, ny);
  const uint32_t zero = 0;
  Mom<R> m = load_mom<PREC>((const T*)a.S + n10, sx, zero, dp);
  R phl = 0, gfx = 0, gfy = 0, gfz = 0, lap = 0, dvt = 0, uox = 0, uoy = 0, uoz = 0;
  R tFx, tFy, tFz;
  if (dp.twocomp) {
    const TA* aux = (const TA*)a.aux;
    const size_t n7 = sbase(7, x, y, z, sx, ny);
    phl = (R)((const TA*)a.phiS)[n];
    const R mg = (R)aux[n7 + 3 * sx];
    gfx = (R)aux[n7] * mg;
    gfy = (R)aux[n7 + sx] * mg;
    gfz = (R)aux[n7 + 2 * sx] * mg;
    lap = (R)((const TA*)a.lap)[n];
    dvt = (R)((const TA*)a.dvt)[n];
    if (dp.dens) {
      const T* uo = (const T*)a.uin;
      uox = C::dec(uo[n3], dp.invscale[1]);
      uoy = C::dec(uo[n3 + sx], dp.invscale[2]);
      uoz = C::dec(uo[n3 + 2 * sx], dp.invscale[3]);
    }
    R rfl[3] = {0, 0, 0};
    if (dp.rep) {
      const TA* rf = (const TA*)a.rep + sbase(6, x, y, z, sx, ny) + 3 * sx;
      rfl[0] = (R)rf[0]; rfl[1] = (R)rf[sx]; rfl[2] = (R)rf[2 * sx];
    }
    moments_stage<true, R>(dp, node_rates<true, R>(dp, phl), phl, gfx, gfy, gfz, lap, dvt, uox, uoy, uoz, m, tFx,
                           tFy, tFz, dp.rep ? rfl : nullptr);
  } else {
    moments_stage<false, R>(dp, node_rates<false, R>(dp, phl), phl, gfx, gfy, gfz, lap, dvt, uox, uoy, uoz, m, tFx,
                            tFy, tFz);
  }
  store_mom<PREC>((T*)a.S + n10, sx, zero, m, dp);
  TA* fr = (TA*)a.frc;
  fr[n3] = (TA)tFx;
  fr[n3 + sx] = (TA)tFy;
  fr[n3 + 2 * sx] = (TA)tFz;
  // the corrected velocity is the next step's u_old (the reference reads it
  // back from this buffer once S and O have swapped)
  const Mom<R> back = load_mom<PREC>((const T*)a.S + n10, sx, zero, dp);
  T* uo = (T*)a.uout;
  uo[n3] = C::enc(back.u, dp.scale[1]);
  uo[n3 + sx] = C::enc(back.v, dp.scale[2]);
  uo[n3 + 2 * sx] = C::enc(back.w, dp.scale[3]);
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
  const int sx = dp.sx;
  const size_t n7 = sbase(7, x, y, z, sx, ny);
  const Phys<R>& ph = dp.ph<R>();
  const TA* aux = (const TA*)a.aux;
  const R t = -ph.sharp_c * (R)((const TA*)a.dvt)[n];
  const R mg = (R)aux[n7 + 3 * sx];
  const R gx = (R)aux[n7] * mg, gy = (R)aux[n7 + sx] * mg, gz = (R)aux[n7 + 2 * sx] * mg;
  const R phv = (R)((const TA*)a.phiS)[n];
  const R lap = (R)((const TA*)a.lap)[n];
  const Mom<R> m = load_mom<PREC>((const T*)a.S + sbase(10, x, y, z, sx, ny), sx, 0u, dp);
  const R o = phv - m.u * (R)0.5 * (gx) - m.v * (R)0.5 * (gy) - m.w * (R)0.5 * (gz) + ph.tau_diff * lap + t;
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
  const int sx = dp.sx;
  const size_t n10 = sbase(10, x, y, z, sx, ny);
  const int nl = dp.lattice == 27 ? 27 : 19;
  const Phys<R>& ph = dp.ph<R>();
  const R cssq = (R)(1.0 / 3.0), invcssq = (R)3, HALF = (R)0.5, ONE = 1, TWO = 2;
  R phl = 0, rhophi = 1;
  if (dp.twocomp) phl = (R)((const TA*)a.phiS)[n];
  if (dp.twocomp && dp.dens) rhophi = ph.rho_r * phl + (ONE - phl) * ph.rho_b;
  const Mom<R> s = load_mom<PREC>((const T*)a.S + n10, sx, 0u, dp);
  Mom<R> o = load_mom<PREC>((const T*)a.O + n10, sx, 0u, dp);
  const R pxx = s.xx - s.u * s.u, pyy = s.yy - s.v * s.v, pzz = s.zz - s.w * s.w;
  const R pxy = s.xy - s.u * s.v, pxz = s.xz - s.u * s.w, pyz = s.yz - s.v * s.w;
  R om;
  if (dp.smag) {
    // boundary:145-160 (division by rho_phi where fused multiplies by its inverse)
    const R visc = dp.twocomp ? (ph.rho_r * ph.visc1 * phl + (ONE - phl) * ph.visc2 * ph.rho_b) / rhophi : ph.visc1;
    const R QQ = pxx * pxx + pyy * pyy + pzz * pzz + TWO * (pxy * pxy + pxz * pxz + pyz * pyz);
    const R t3 = (R)3.0 * visc;
    om = (R)0.5 + (R)(1.0 / 6.0) * ((R)3.0 * visc + sqrt(t3 * t3 + (R)(0.053f * 18.0f) * sqrt((R)2.0 * QQ) / rhophi));
    om = (R)1.0 / om;
  } else if (dp.twocomp) {
    const R visc = (ph.rho_r * ph.visc1 * phl + (ONE - phl) * ph.visc2 * ph.rho_b) / rhophi;
    om = (visc * invcssq + (R)0.5);
    om = (R)1.0 / om;
  } else {
    om = ph.omega;
  }
  for (int l = 1; l < nl; ++l) {
    // upstream node x - c_l (= x + c_opp(l))
    if (isf_at(a.isfh, x - cEX[l], y - cEY[l], z - cEZ[l], nx, ny, nz) != 0) continue;
    const R ex = (R)cEX[l], ey = (R)cEY[l], ez = (R)cEZ[l];
    const int d2 = cEX[l] * cEX[l] + cEY[l] * cEY[l] + cEZ[l] * cEZ[l];
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
  store_mom<PREC>((T*)a.O + n10, sx, 0u, o, dp);
}

// ---------------------------------------------------------------------------
// REPULSIVE_FLUX (lb_cuda_repulsive_module.f90).  Neighbours up to three
// nodes away wrap across periodic faces (the reference's 3-deep halo,
// mpi_module.f90:163-165) and are skipped across a wall face.
__device__ __forceinline__ bool wrap_win(int& x, int& y, int& z, const DevParams& dp) {
  auto w = [](int& a, int n, int per) {
    if (a >= 0 && a < n) return true;
    if (!per) return false;
    a = ((a % n) + n) % n;
    return true;
  };
  return w(x, dp.nx, dp.pbc[0]) && w(y, dp.ny, dp.pbc[1]) && w(z, dp.nz, dp.pbc[2]);
}

// thinfilm_scan_mark_kernel (repulsive:19-150): for an interfacial node, the
// nearest facing interfacial node in the 7^3 window (ties: larger facing factor).
template <int PREC>
__global__ void __launch_bounds__(BS) k_thinfilm_scan(StageArgs a) {
  typedef typename Codec<PREC>::R R;
  typedef typename Aux<PREC>::T TA;
  const DevParams& dp = a.dp;
  int x, y, z;
  if (!node_of((size_t)blockIdx.x * BS + threadIdx.x, dp, x, y, z)) return;
  const int nx = dp.nx, ny = dp.ny, nz = dp.nz;
  const int8_t isf = isf_at(a.isfh, x, y, z, nx, ny, nz);
  if (isf != 1 && isf != -1) return;
  const size_t n = nidx(x, y, z, nx, ny);
  const int sx = dp.sx;
  const size_t n6 = sbase(6, x, y, z, sx, ny), n7 = sbase(7, x, y, z, sx, ny);
  const Phys<R>& ph = dp.ph<R>();
  const TA* phi = (const TA*)a.phiS;
  const TA* aux = (const TA*)a.aux;
  TA* loc = (TA*)a.rep;
  a.rep_mask[n] = 0;
  loc[n6] = (TA)0; loc[n6 + sx] = (TA)0; loc[n6 + 2 * sx] = (TA)0;
  R q = fmin(fmax((R)phi[n], (R)0), (R)1);
  q = q * ((R)1 - q);
  if (q < ph.q_th) return;
  const R nix = (R)aux[n7], niy = (R)aux[n7 + sx], niz = (R)aux[n7 + 2 * sx];
  R best_r2 = (R)1.0e300, best_face = (R)-1;
  int bi = 0, bj = 0, bk = 0;
  bool found = false;
  for (int di = -3; di <= 3; ++di)
    for (int dj = -3; dj <= 3; ++dj)
      for (int dk = -3; dk <= 3; ++dk) {
        if (di == 0 && dj == 0 && dk == 0) continue;
        int xi = x + di, yj = y + dj, zk = z + dk;
        if (!wrap_win(xi, yj, zk, dp)) continue;
        const int8_t v = isf_at(a.isfh, xi, yj, zk, nx, ny, nz);
        if (v != 1 && v != -1) continue;
        const R r2 = (R)(di * di + dj * dj + dk * dk);
        const size_t m = nidx(xi, yj, zk, nx, ny), m7 = sbase(7, xi, yj, zk, sx, ny);
        R qn = fmin(fmax((R)phi[m], (R)0), (R)1);
        qn = qn * ((R)1 - qn);
        if (qn < ph.q_th || fabs(qn - q) > (R)0.1 * fmax(q, (R)1.0e-12)) continue;
        const R dotn = nix * (R)aux[m7] + niy * (R)aux[m7 + sx] + niz * (R)aux[m7 + 2 * sx];
        if (dotn > ph.cos_opp) continue;
        const R face = (R)0.5 * ((R)1 - dotn);
        if (r2 < best_r2 - (R)1.0e-14) {
          best_r2 = r2; best_face = face; bi = x + di; bj = y + dj; bk = z + dk; found = true;
        } else if (fabs(r2 - best_r2) <= (R)1.0e-14 && face > best_face) {
          best_face = face; bi = x + di; bj = y + dj; bk = z + dk; found = true;
        }
      }
  if (found) {
    // the reference stores 1-based local indices
    loc[n6] = (TA)(bi + 1); loc[n6 + sx] = (TA)(bj + 1); loc[n6 + 2 * sx] = (TA)(bk + 1);
    a.rep_mask[n] = 1;
  }
}

// repulsive_flux_normal_kernel (repulsive:152-289)
template <int PREC>
__global__ void __launch_bounds__(BS) k_repulsive_flux(StageArgs a) {
  typedef typename Codec<PREC>::R R;
  typedef typename Aux<PREC>::T TA;
  const DevParams& dp = a.dp;
  int x, y, z;
  if (!node_of((size_t)blockIdx.x * BS + threadIdx.x, dp, x, y, z)) return;
  const int nx = dp.nx, ny = dp.ny, nz = dp.nz;
  const int8_t isf = isf_at(a.isfh, x, y, z, nx, ny, nz);
  if (isf != 1 && isf != -1) return;
  const size_t n = nidx(x, y, z, nx, ny);
  const int sx = dp.sx;
  const size_t n6 = sbase(6, x, y, z, sx, ny), n7 = sbase(7, x, y, z, sx, ny);
  const Phys<R>& ph = dp.ph<R>();
  const TA* phi = (const TA*)a.phiS;
  const TA* aux = (const TA*)a.aux;
  TA* loc = (TA*)a.rep;
  TA* flux = loc + 3 * sx;
  flux[n6] = (TA)0; flux[n6 + sx] = (TA)0; flux[n6 + 2 * sx] = (TA)0;
  if (a.rep_mask[n] != 1) return;
  const R eps = (R)1.0e-9;
  const R phl = (R)phi[n];
  const R q1 = phl * ((R)1 - phl);
  if (q1 <= eps) return;
  const int pi = (int)(R)loc[n6] - 1, pj = (int)(R)loc[n6 + sx] - 1, pk = (int)(R)loc[n6 + 2 * sx] - 1;
  R dx = (R)(pi - x), dy = (R)(pj - y), dz = (R)(pk - z);
  const R r = sqrt(dx * dx + dy * dy + dz * dz);
  if (r <= eps) return;
  const R rinv = (R)1.0 / r;
  dx = dx * rinv; dy = dy * rinv; dz = dz * rinv;
  int xi = pi, yj = pj, zk = pk;
  if (!wrap_win(xi, yj, zk, dp)) return;
  const size_t m = nidx(xi, yj, zk, nx, ny), m7 = sbase(7, xi, yj, zk, sx, ny);
  const R nx1 = (R)aux[n7], ny1 = (R)aux[n7 + sx], nz1 = (R)aux[n7 + 2 * sx];
  const R nx2 = (R)aux[m7], ny2 = (R)aux[m7 + sx], nz2 = (R)aux[m7 + 2 * sx];
  const R face = fmax((R)0, -(nx1 * nx2 + ny1 * ny2 + nz1 * nz2));
  if (face <= eps) return;
  R nsx = nx1 - nx2, nsy = ny1 - ny2, nsz = nz1 - nz2;
  const R nsmag = sqrt(nsx * nsx + nsy * nsy + nsz * nsz);
  if (nsmag <= eps) {
    nsx = dx; nsy = dy; nsz = dz;
  } else {
    nsx = nsx / nsmag; nsy = nsy / nsmag; nsz = nsz / nsmag;
  }
  if (dx * nsx + dy * nsy + dz * nsz < (R)0) { nsx = -nsx; nsy = -nsy; nsz = -nsz; }
  const R phl2 = (R)phi[m];
  const R q2 = phl2 * ((R)1 - phl2);
  const R qpair = (R)0.5 * (q1 + q2);
  const R qcl = fmin(fmax(qpair, eps), (R)0.25 - eps);
  const R arg = (R)1.0 / ((R)2.0 * sqrt(qcl));
  if (arg <= (R)1.0) return;
  const R ach = log(arg + sqrt(arg * arg - (R)1.0));
  const R Wfilm = ph.width * ach;
  if (Wfilm <= (R)0) return;
  const R w2 = Wfilm * Wfilm;
  const R wdth = (R)1.0 / ((R)1.0 + w2 * w2);
  const R fx = ph.a_rep * wdth * qcl * face * nsx;
  const R fy = ph.a_rep * wdth * qcl * face * nsy;
  const R fz = ph.a_rep * wdth * qcl * face * nsz;
  const R cap = (R)1.5 * (fabs(fx) + fabs(fy) + fabs(fz));
  const R scales = fmin((R)1.0, phl / fmax(cap, (R)1.0e-9));
  flux[n6] = (TA)(fx * scales); flux[n6 + sx] = (TA)(fy * scales); flux[n6 + 2 * sx] = (TA)(fz * scales);
}

// ---------------------------------------------------------------------------
// Open boundaries (openbc = 1, internal obstacles; boundary_cds_module.f90).
// Face f = 2*axis + side of every non-periodic axis.  Its wall is the solid
// layer at coordinate 0 (side 0) or n-1 (side 1) along the axis; the
// moments closure works on the first fluid layer, coordinate 1 or n-2.  The
// reference's face loops skip the face's edges: the two other coordinates
// run over 1 .. n-2 (gj > 1 .and. gj < ly, 0-based here).
__device__ __forceinline__ bool open_face_node(const DevParams& dp, int f, int x, int y, int z, int depth) {
  const int ax = f >> 1, side = f & 1;
  if (dp.pbc[ax]) return false;
  const int c[3] = {x, y, z}, n[3] = {dp.nx, dp.ny, dp.nz};
  if (c[ax] != (side ? n[ax] - 1 - depth : depth)) return false;
  for (int b = 0; b < 3; ++b)
    if (b != ax && (c[b] < 1 || c[b] > n[b] - 2)) return false;
  return true;
}

// The open-face kernels run over the face layers only: blockIdx.y is the
// face f (axis f >> 1, side f & 1) and the threads of a face cover its layer
// at the given depth from the wall, the rectangle [1, n - 2] on the two other
// axes (the nodes open_face_node accepts).  A node on several faces (depth
// 1: where two open faces meet) is taken by the thread of its first face.
__host__ __device__ inline size_t open_face_area(const DevParams& dp) {
  const int n[3] = {dp.nx, dp.ny, dp.nz};
  size_t m = 0;
  for (int ax = 0; ax < 3; ++ax) {
    if (dp.pbc[ax]) continue;
    const int w1 = n[ax == 0 ? 1 : 0] - 2, w2 = n[ax == 2 ? 1 : 2] - 2;
    if (w1 > 0 && w2 > 0 && (size_t)w1 * w2 > m) m = (size_t)w1 * w2;
  }
  return m;
}

__device__ __forceinline__ bool open_face_thread(const DevParams& dp, int depth, int& x, int& y, int& z) {
  const int f = blockIdx.y, ax = f >> 1, side = f & 1;
  if (dp.pbc[ax]) return false;
  const int n[3] = {dp.nx, dp.ny, dp.nz};
  const int b1 = ax == 0 ? 1 : 0, b2 = ax == 2 ? 1 : 2;
  const int w1 = n[b1] - 2, w2 = n[b2] - 2;
  if (w1 <= 0 || w2 <= 0) return false;
  const size_t t = (size_t)blockIdx.x * BS + threadIdx.x;
  if (t >= (size_t)w1 * w2) return false;
  int c[3];
  c[ax] = side ? n[ax] - 1 - depth : depth;
  if (c[ax] < 0 || c[ax] >= n[ax]) return false;
  c[b1] = 1 + (int)(t % (size_t)w1);
  c[b2] = 1 + (int)(t / (size_t)w1);
  x = c[0]; y = c[1]; z = c[2];
  for (int g = 0; g < f; ++g)
    if (open_face_node(dp, g, x, y, z, depth)) return false;
  return true;
}

// Phase field (boundary_cds_module.f90:1353-1664): a wall node of an open
// face takes phi of the node next to it inside the domain.  The faces do not
// overlap (edges are skipped) and no node read is written.
template <int PREC>
__global__ void __launch_bounds__(BS) k_open_phi(StageArgs a) {
  typedef typename Aux<PREC>::T TA;
  const DevParams& dp = a.dp;
  int x, y, z;
  if (!open_face_thread(dp, 0, x, y, z)) return;
  for (int f = 0; f < 6; ++f) {
    if (!open_face_node(dp, f, x, y, z, 0)) continue;
    const int ax = f >> 1, d = (f & 1) ? -1 : 1;
    TA* phi = (TA*)a.phiS;
    phi[nidx(x, y, z, dp.nx, dp.ny)] =
        phi[nidx(x + (ax == 0) * d, y + (ax == 1) * d, z + (ax == 2) * d, dp.nx, dp.ny)];
  }
}

// Moments (boundary_cds_module.f90:57-1311), after LB_int_boundary, on the
// first fluid layer (isfluid -1) of each open face: for every population that
// a node receives from a solid node, take out the near-wall closure's
// population, w_l P + (1 - omega) fneq_l with the non-equilibrium stress
// Pi - cs^2 P - u u of the node, and put in the regularised population of the
// face state -- P and u of the node, or the face's pressure (type 1) or
// velocity (type 0) in their place -- with the same stress and, under
// HIGHORDER, the E_k / N_k terms.  The outgoing second moments of the removed
// population are taken without the cs^2 shift (fpost c_a c_b), those of the
// added one with it, as the reference writes them.  S is the post-moments
// state, O the output of the fused kernel and LB_int_boundary; a node on two
// faces (next to an edge of the layer) is corrected for both, in face order.
template <int PREC>
__global__ void __launch_bounds__(BS) k_open_h(StageArgs a) {
  typedef Codec<PREC> C;
  typedef typename C::R R;
  typedef typename C::T T;
  typedef typename Aux<PREC>::T TA;
  const DevParams& dp = a.dp;
  int x, y, z;
  if (!open_face_thread(dp, 1, x, y, z)) return;
  const int nx = dp.nx, ny = dp.ny, nz = dp.nz;
  if (isf_at(a.isfh, x, y, z, nx, ny, nz) != -1) return;
  const size_t n = nidx(x, y, z, nx, ny);
  const int sx = dp.sx;
  const size_t n10 = sbase(10, x, y, z, sx, ny);
  const Phys<R>& ph = dp.ph<R>();
  const R cssq = (R)(1.0 / 3.0), ONE = 1;
  const Mom<R> s = load_mom<PREC>((const T*)a.S + n10, sx, 0u, dp);
  Mom<R> o = load_mom<PREC>((const T*)a.O + n10, sx, 0u, dp);
  R om;
  if (dp.twocomp) {
    const R phl = (R)((const TA*)a.phiS)[n];
    const R rhophi = dp.dens ? ph.rho_r * phl + (ONE - phl) * ph.rho_b : ONE;
    const R visc = (ph.rho_r * ph.visc1 * phl + (ONE - phl) * ph.visc2 * ph.rho_b) / rhophi;
    om = ONE / (visc * (R)3 + (R)0.5);
  } else {
    om = ph.omega;
  }
  NodeIn<R> sub{}, add{};
  sub.P = s.P;
  sub.xx = s.xx - cssq * s.P - s.u * s.u; sub.yy = s.yy - cssq * s.P - s.v * s.v;
  sub.zz = s.zz - cssq * s.P - s.w * s.w;
  sub.xy = s.xy - s.u * s.v; sub.xz = s.xz - s.u * s.w; sub.yz = s.yz - s.v * s.w;
  sub.om = om;
  R fs[27], fa[27];
  if (dp.lattice == 27) collide<27, false, R>(sub, fs);
  else collide<19, false, R>(sub, fs);
  for (int f = 0; f < 6; ++f) {
    if (!open_face_node(dp, f, x, y, z, 1)) continue;
    add = sub;
    add.u = s.u; add.v = s.v; add.w = s.w;
    const int type = a.obc_type[f];
    if (type == 1) {
      add.P = (R)a.obc_val[f][0];
    } else if (type != 2) {
      add.u = (R)a.obc_val[f][1]; add.v = (R)a.obc_val[f][2]; add.w = (R)a.obc_val[f][3];
    }
    if (dp.lattice == 19) collide<19, false, R>(add, fa);
    else if (dp.highorder) collide<27, true, R>(add, fa);
    else collide<27, false, R>(add, fa);
    const int nl = dp.lattice == 27 ? 27 : 19;
    for (int l = 1; l < nl; ++l) {
      if (isf_at(a.isfh, x - cEX[l], y - cEY[l], z - cEZ[l], nx, ny, nz) != 0) continue;
      const R ex = (R)cEX[l], ey = (R)cEY[l], ez = (R)cEZ[l];
      const int ci = cidx(cEX[l], cEY[l], cEZ[l]);
      const R fr = fs[ci];
      o.P = o.P - fr;
      o.u = o.u - fr * ex; o.v = o.v - fr * ey; o.w = o.w - fr * ez;
      o.xx = o.xx - fr * ex * ex; o.yy = o.yy - fr * ey * ey; o.zz = o.zz - fr * ez * ez;
      o.xy = o.xy - fr * ex * ey; o.xz = o.xz - fr * ex * ez; o.yz = o.yz - fr * ey * ez;
      const R fp = fa[ci];
      o.P = o.P + fp;
      o.u = o.u + fp * ex; o.v = o.v + fp * ey; o.w = o.w + fp * ez;
      o.xx = o.xx + fp * (ex * ex - cssq); o.yy = o.yy + fp * (ey * ey - cssq); o.zz = o.zz + fp * (ez * ez - cssq);
      o.xy = o.xy + fp * ex * ey; o.xz = o.xz + fp * ex * ez; o.yz = o.yz + fp * ey * ez;
    }
  }
  store_mom<PREC>((T*)a.O + n10, sx, 0u, o, dp);
}

// ---------------------------------------------------------------------------
template <int PREC>
static hipError_t stage_t(int stage, const StageArgs& a, hipStream_t s) {
  const size_t n = a.dp.nn;
  const dim3 grid((unsigned)((n + BS - 1) / BS)), block(BS);
  switch (stage) {
    case 1: hipLaunchKernelGGL(k_norm<PREC>, grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_phi_bnd<PREC>, grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL(k_div<PREC>, grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL(k_moments<PREC>, grid, block, 0, s, a); break;
    case 6: hipLaunchKernelGGL(k_update_phi<PREC>, grid, block, 0, s, a); break;
    case 7: hipLaunchKernelGGL(k_lb_int_bnd<PREC>, grid, block, 0, s, a); break;
    case 8: hipLaunchKernelGGL(k_thinfilm_scan<PREC>, grid, block, 0, s, a); break;
    case 9: hipLaunchKernelGGL(k_repulsive_flux<PREC>, grid, block, 0, s, a); break;
    case 11:
    case 12: {
      const size_t m = open_face_area(a.dp);
      if (m == 0) break;  // no open face has nodes
      const dim3 fgrid((unsigned)((m + BS - 1) / BS), 6);
      if (stage == 11) hipLaunchKernelGGL(k_open_phi<PREC>, fgrid, block, 0, s, a);
      else hipLaunchKernelGGL(k_open_h<PREC>, fgrid, block, 0, s, a);
      break;
    }
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

template <int PREC>
static hipError_t prep_t(const StageArgs& a, int tz, hipStream_t s) {
  if constexpr (sizeof(typename Aux<PREC>::T) == 4) {
    if (a.div_only) hipLaunchKernelGGL((k_phase_prep_z<PREC, true>), dim3(prep_tiles_z(a.dp, tz)), dim3(QNT), 0, s, a, tz);
    else hipLaunchKernelGGL((k_phase_prep_z<PREC, false>), dim3(prep_tiles_z(a.dp, tz)), dim3(QNT), 0, s, a, tz);
  } else {
    const dim3 grid(((a.dp.nx + PTX - 1) / PTX) * ((a.dp.ny + PTY - 1) / PTY) * ((a.dp.nz + tz - 1) / tz));
    hipLaunchKernelGGL(k_phase_prep<PREC>, grid, dim3(PNT), 0, s, a, tz);
  }
  return hipGetLastError();
}

int prep_tile_count(const StageArgs& a, int prec, int tz) {
  if (prec == PREC_FP64) return -1;  // the FP64 prep is not batched
  return prep_tiles_z(a.dp, tz);
}

hipError_t launch_phase_prep_batch(const PrepBatchEntry* dev_tab, int ndom, int total_tiles, int prec, hipStream_t s) {
  const dim3 grid(total_tiles), block(QNT);
  // batches hold two-component domains without obstacles: div theta only
  // exactly where fused_grad_inline() holds, i.e. for every precision but MIXED
  switch (prec) {
    case PREC_FP32: hipLaunchKernelGGL((k_phase_prep_z_batch<PREC_FP32, true>), grid, block, 0, s, dev_tab, ndom, total_tiles); break;
    case PREC_MIXED: hipLaunchKernelGGL((k_phase_prep_z_batch<PREC_MIXED, false>), grid, block, 0, s, dev_tab, ndom, total_tiles); break;
    case PREC_FP16S: hipLaunchKernelGGL((k_phase_prep_z_batch<PREC_FP16S, true>), grid, block, 0, s, dev_tab, ndom, total_tiles); break;
    case PREC_BF16S: hipLaunchKernelGGL((k_phase_prep_z_batch<PREC_BF16S, true>), grid, block, 0, s, dev_tab, ndom, total_tiles); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_phase_prep(const StageArgs& a, int prec, int tz, hipStream_t s) {
  switch (prec) {
    case PREC_FP64: return prep_t<PREC_FP64>(a, tz, s);
    case PREC_FP32: return prep_t<PREC_FP32>(a, tz, s);
    case PREC_MIXED: return prep_t<PREC_MIXED>(a, tz, s);
    case PREC_FP16S: return prep_t<PREC_FP16S>(a, tz, s);
    case PREC_BF16S: return prep_t<PREC_BF16S>(a, tz, s);
  }
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// encode / decode
template <int PREC>
__global__ void __launch_bounds__(BS) k_encode(StageArgs a, const float* raw) {
  typedef Codec<PREC> C;
  typedef typename C::R R;
  typedef typename C::T T;
  int x, y, z;
  if (!node_of((size_t)blockIdx.x * BS + threadIdx.x, a.dp, x, y, z)) return;
  const int sx = a.dp.sx;
  const size_t n = sbase(10, x, y, z, sx, a.dp.ny);  // raw is a 10-field FP32 set
  Mom<R> m;
  m.P = raw[n]; m.u = raw[n + sx]; m.v = raw[n + 2 * sx]; m.w = raw[n + 3 * sx];
  m.xx = raw[n + 4 * sx]; m.yy = raw[n + 5 * sx]; m.zz = raw[n + 6 * sx];
  m.xy = raw[n + 7 * sx]; m.xz = raw[n + 8 * sx]; m.yz = raw[n + 9 * sx];
  store_mom<PREC>((T*)a.S + n, sx, 0u, m, a.dp);
}

hipError_t launch_encode_moments(void* S, const float* raw, const DevParams& dp, int prec, hipStream_t s) {
  StageArgs a{};
  a.S = S;
  a.dp = dp;
  const dim3 grid((unsigned)((dp.nn + BS - 1) / BS)), block(BS);
  if (prec == PREC_FP16S)
    hipLaunchKernelGGL(k_encode<PREC_FP16S>, grid, block, 0, s, a, raw);
  else if (prec == PREC_BF16S)
    hipLaunchKernelGGL(k_encode<PREC_BF16S>, grid, block, 0, s, a, raw);
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
  int x, y, z;
  if (!node_of(n, a.dp, x, y, z)) return;
  const Mom<R> m = load_mom<PREC>((const T*)a.S + sbase(10, x, y, z, a.dp.sx, a.dp.ny), a.dp.sx, 0u, a.dp);
  const R v[10] = {m.P, m.u, m.v, m.w, m.xx, m.yy, m.zz, m.xy, m.xz, m.yz};
  R r = 0;
#pragma unroll
  for (int i = 0; i < 10; ++i)
    if (i == comp) r = v[i];
  out[n] = (O)r;
}

template <int PREC>
static hipError_t decode_t(const StageArgs& a, int comp, void* out, int dtype, hipStream_t s) {
  const dim3 grid((unsigned)((a.dp.nn + BS - 1) / BS)), block(BS);
  if (dtype == 0)
    hipLaunchKernelGGL((k_decode<PREC, double>), grid, block, 0, s, a, comp, (double*)out);
  else
    hipLaunchKernelGGL((k_decode<PREC, float>), grid, block, 0, s, a, comp, (float*)out);
  return hipGetLastError();
}

hipError_t launch_decode_field(const void* S, int comp, void* out, int out_dtype, const DevParams& dp, int prec,
                               hipStream_t s) {
  StageArgs a{};
  a.S = const_cast<void*>(S);
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
// element index of host-order node i in a field of layout l
__device__ __forceinline__ size_t lay_at(FieldLay l, size_t i, int nx, int sx) {
  if (l.F == 0) return i;
  const size_t row = i / (size_t)nx;
  return (row * l.F + l.f) * (size_t)sx + (i - row * (size_t)nx);
}
__global__ void __launch_bounds__(BS) k_convert(const void* src, int sk, FieldLay sl, void* dst, int dk, FieldLay dl,
                                                size_t n, int nx, int sx, float inv, float sc) {
  const size_t i = (size_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  kstore(dst, dk, lay_at(dl, i, nx, sx), kload(src, sk, lay_at(sl, i, nx, sx), inv), sc);
}

hipError_t launch_convert_in(const void* src, int src_dtype, void* dst, int dst_kind, FieldLay dl, const DevParams& dp,
                             float scale, hipStream_t s) {
  const dim3 grid((unsigned)((dp.nn + BS - 1) / BS)), block(BS);
  hipLaunchKernelGGL(k_convert, grid, block, 0, s, src, src_dtype, FieldLay{0, 0}, dst, dst_kind, dl, dp.nn, dp.nx,
                     dp.sx, 1.0f, scale);
  return hipGetLastError();
}
hipError_t launch_convert_out(const void* src, int src_kind, FieldLay sl, void* dst, int dst_dtype, const DevParams& dp,
                              float invscale, hipStream_t s) {
  const dim3 grid((unsigned)((dp.nn + BS - 1) / BS)), block(BS);
  hipLaunchKernelGGL(k_convert, grid, block, 0, s, src, src_kind, sl, dst, dst_dtype, FieldLay{0, 0}, dp.nn, dp.nx,
                     dp.sx, invscale, 1.0f);
  return hipGetLastError();
}

template <class E>
__global__ void __launch_bounds__(BS) k_set_copy(E* dst, FieldLay dl, const E* src, FieldLay sl, int nf, size_t n,
                                                 int nx, int sx) {
  const size_t i = (size_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  for (int k = 0; k < nf; ++k)
    dst[lay_at(FieldLay{dl.F, dl.f + k}, i, nx, sx)] = src[lay_at(FieldLay{sl.F, sl.f + k}, i, nx, sx)];
}
hipError_t launch_set_copy(void* dst, FieldLay dl, const void* src, FieldLay sl, int nf, int esz, const DevParams& dp,
                           hipStream_t s) {
  const dim3 grid((unsigned)((dp.nn + BS - 1) / BS)), block(BS);
  switch (esz) {
    case 2: hipLaunchKernelGGL(k_set_copy<uint16_t>, grid, block, 0, s, (uint16_t*)dst, dl, (const uint16_t*)src, sl, nf, dp.nn, dp.nx, dp.sx); break;
    case 4: hipLaunchKernelGGL(k_set_copy<uint32_t>, grid, block, 0, s, (uint32_t*)dst, dl, (const uint32_t*)src, sl, nf, dp.nn, dp.nx, dp.sx); break;
    case 8: hipLaunchKernelGGL(k_set_copy<uint64_t>, grid, block, 0, s, (uint64_t*)dst, dl, (const uint64_t*)src, sl, nf, dp.nn, dp.nx, dp.sx); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

__global__ void k_tick(uint32_t* ctr, int n) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d < n) ctr[d] += 1u;
}

hipError_t launch_tick(uint32_t* ctr, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_tick, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ctr, n);
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
// BCPHIFLUX: the global phase-mass correction of phi_sum_count_kernel and
// apply_lagrangian_phi_kernel (lb_cuda_boundary_module.f90:461-566), driven as
// boundary_cds_module.f90:1864-1884 drives them: the sum of phi over the
// |isfluid| == 1 nodes and the count of those with 0.5 < phi < 0.9, then
// corr = (sum_ini - sum) / count added to the counted nodes.  The sum is
// taken in a fixed order (block partials, then one thread), the correction
// is formed on the device, so no host round trip sits in the step.
// lag[0] = sum at initialisation, lag[1] = last sum, lag[2] = last count,
// lag[3] = last correction.
template <int PREC>
__global__ void __launch_bounds__(BS) k_phi_lag_sum(StageArgs a, double* partial) {
  typedef typename Aux<PREC>::T TA;
  typedef typename Codec<PREC>::T T;
  __shared__ double sh[3][BS];
  const DevParams& dp = a.dp;
  double sum = 0, cnt = 0, flux = 0;
  for (size_t t = (size_t)blockIdx.x * BS + threadIdx.x; t < dp.nn; t += (size_t)gridDim.x * BS) {
    int x, y, z;
    node_of(t, dp, x, y, z);
    const int8_t isf = isf_at(a.isfh, x, y, z, dp.nx, dp.ny, dp.nz);
    if (isf != 1 && isf != -1) continue;
    const double ph = (double)((const TA*)a.phiS)[t];
    sum += ph;
    if (ph > 0.5 && ph < 0.9) cnt += 1.0;
    if (a.obc && isf == -1) {
      // with open boundaries, the phase carried through each open face's
      // first fluid layer (boundary_cds_module.f90:1667-1862): + u phi on
      // the low face, - u phi on the high face, u the velocity along the
      // face's axis in S; no edge is skipped here
      const int c[3] = {x, y, z}, nn[3] = {dp.nx, dp.ny, dp.nz};
      for (int ax = 0; ax < 3; ++ax) {
        if (dp.pbc[ax]) continue;
        const bool lo = c[ax] == 1, hi = c[ax] == nn[ax] - 2;
        if (!lo && !hi) continue;
        const T* S = (const T*)a.S + sbase(10, x, y, z, dp.sx, dp.ny);
        const double ua = (double)Codec<PREC>::dec(S[(size_t)(1 + ax) * dp.sx], dp.invscale[1 + ax]);
        if (lo) flux += ua * ph;
        if (hi) flux -= ua * ph;
      }
    }
  }
  sh[0][threadIdx.x] = sum;
  sh[1][threadIdx.x] = cnt;
  sh[2][threadIdx.x] = flux;
  __syncthreads();
  for (int st = BS / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + st];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + st];
      sh[2][threadIdx.x] += sh[2][threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x < 3) partial[(size_t)blockIdx.x * 3 + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void k_phi_lag_final(const double* partial, int nb, double* lag, int init) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double sum = 0, cnt = 0, change = 0;
  for (int b = 0; b < nb; ++b) {
    sum += partial[3 * b];
    cnt += partial[3 * b + 1];
    change += partial[3 * b + 2];
  }
  if (init) {
    lag[0] = sum;  // global_phi_sum_ini (init_conditions_module.f90:469-490)
    lag[1] = sum;
    lag[2] = cnt;
    lag[3] = 0.0;
    return;
  }
  lag[1] = sum;
  lag[2] = cnt;
  // corr = dphi / real(global_count): the count is converted to a default
  // (single precision) real, as the reference's real() without a kind does.
  // No counted node (an empty interface band) gives no correction.
  // With open boundaries the target adds this step's flux through the open
  // faces (dphi = (global_phi_sum_ini + global_phi_change) - global_phi_sum,
  // boundary_cds_module.f90:1864-1870; the flux is reset every step).
  lag[3] = cnt > 0 ? ((lag[0] + change) - sum) / (double)(float)cnt : 0.0;
}

template <int PREC>
__global__ void __launch_bounds__(BS) k_phi_lag_apply(StageArgs a, const double* lag) {
  typedef typename Aux<PREC>::T TA;
  const DevParams& dp = a.dp;
  int x, y, z;
  const size_t t = (size_t)blockIdx.x * BS + threadIdx.x;
  if (!node_of(t, dp, x, y, z)) return;
  const int8_t isf = isf_at(a.isfh, x, y, z, dp.nx, dp.ny, dp.nz);
  if (isf != 1 && isf != -1) return;
  TA* phi = (TA*)a.phiS;
  const double ph = (double)phi[t];
  if (ph > 0.5 && ph < 0.9) phi[t] = (TA)(ph + lag[3]);
}

template <int PREC>
static hipError_t lag_t(const StageArgs& a, double* partial, int nb, double* lag, bool init, hipStream_t s) {
  hipLaunchKernelGGL(k_phi_lag_sum<PREC>, dim3(nb), dim3(BS), 0, s, a, partial);
  hipLaunchKernelGGL(k_phi_lag_final, dim3(1), dim3(64), 0, s, (const double*)partial, nb, lag, init ? 1 : 0);
  if (!init)
    hipLaunchKernelGGL(k_phi_lag_apply<PREC>, dim3((unsigned)((a.dp.nn + BS - 1) / BS)), dim3(BS), 0, s, a,
                       (const double*)lag);
  return hipGetLastError();
}

hipError_t launch_phi_lagrange(const StageArgs& a, int prec, double* partial, int nblocks, double* lag, bool init,
                               hipStream_t s) {
  switch (prec) {
    case PREC_FP64: return lag_t<PREC_FP64>(a, partial, nblocks, lag, init, s);
    case PREC_FP32: return lag_t<PREC_FP32>(a, partial, nblocks, lag, init, s);
    case PREC_MIXED: return lag_t<PREC_MIXED>(a, partial, nblocks, lag, init, s);
    case PREC_FP16S: return lag_t<PREC_FP16S>(a, partial, nblocks, lag, init, s);
    case PREC_BF16S: return lag_t<PREC_BF16S>(a, partial, nblocks, lag, init, s);
  }
  return hipErrorInvalidValue;
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
  for (size_t t = (size_t)blockIdx.x * BS + threadIdx.x; t < dp.nn; t += (size_t)gridDim.x * BS) {
    int x, y, z;
    node_of(t, dp, x, y, z);
    const int8_t isf = isf_at(a.isfh, x, y, z, dp.nx, dp.ny, dp.nz);
    if (isf != 1 && isf != -1) continue;
    const Mom<R> m = load_mom<PREC>((const T*)a.S + sbase(10, x, y, z, dp.sx, dp.ny), dp.sx, 0u, dp);
    double ph = 0, rho = 1;
    if (dp.twocomp) {
      ph = (double)((const TA*)a.phiS)[t];
      rho = dp.dens ? dp.pd.rho_r * ph + (1.0 - ph) * dp.pd.rho_b : 1.0;
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

hipError_t launch_reduce(const void* S, const void* phi, const int8_t* isfh, const DevParams& dp, int prec,
                         double* partial, int nblocks, double* out8, hipStream_t s) {
  StageArgs a{};
  a.S = const_cast<void*>(S);
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

// ---------------------------------------------------------------------------
// Reductions that allocate nothing: the per-block partials and the eight sums
// live in `scratch`, a file-scope device array that each GPU holds until the
// process ends (64 KiB + 64 B).  Host code never passes the array's name to a
// kernel or a copy: on the host it is only a stand-in for the symbol.  The
// device address comes from hipGetSymbolAddress, once per device (with that
// device current), and is kept per device: every GPU has its own copy.
// Calls that overlap on one GPU share the array and race.
static __device__ __align__(256) double scratch[REDUCE_SCRATCH_BLOCKS * 8 + 8];

static std::mutex g_scratch_mu;
static double* g_scratch_ptr[REDUCE_SCRATCH_MAX_DEVICES] = {};

hipError_t reduce_scratch_ptr(int device, double** out) {
  *out = nullptr;
  if (device < 0 || device >= REDUCE_SCRATCH_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) {
      const hipError_t e2 = hipSetDevice(cur);
      if (e == hipSuccess) e = e2;
    }
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = (double*)p;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

// The same kernels as launch_reduce, on min(REDUCE_SCRATCH_BLOCKS, ceil(nn /
// BS)) blocks.  A node's (block, thread) and every partial sum are those of a
// REDUCE_SCRATCH_BLOCKS-block launch (its extra blocks hold exact zeros), so
// the sums equal lbf_reduce_fp64's bit for bit.  The grid-stride loop covers
// any domain size, so the scratch bounds the block count, not the domain.
hipError_t launch_reduce_scratch(const void* S, const void* phi, const int8_t* isfh, const DevParams& dp, int prec,
                                 double* scratch_dev, hipStream_t s) {
  const size_t want = (dp.nn + BS - 1) / BS;
  const int nb = (int)(want < (size_t)REDUCE_SCRATCH_BLOCKS ? (want ? want : 1) : (size_t)REDUCE_SCRATCH_BLOCKS);
  return launch_reduce(S, phi, isfh, dp, prec, scratch_dev, nb, scratch_dev + REDUCE_SCRATCH_SUMS, s);
}

}  // namespace lbf
