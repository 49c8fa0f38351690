// lbf_kernels.h -- launch interfaces between the C-ABI layer and the kernels.
#pragma once
#include "lbf_common.h"

namespace lbf {

// Field sets are row-interleaved (sbase() in lbf_common.h); single fields
// (phi, lap, div theta, masks) are plain x-fastest arrays.
struct FusedArgs {
  const void* min;   // S moments: 10-field set
  void* mout;        // O moments
  const void* uin;   // previous corrected velocity: 3-field set (2c DENSRATIO)
  void* uout;        // this step's corrected velocity
  const void* phin;  // phase field S
  void* phiout;      // phase field O
  const void* pre;   // 5-field set: grad phi (3, as n*|g|), lap phi, div(theta n) (inline)
  const void* frc;   // forces: 3-field set (stage path)
  const int8_t* isfh;  // isfluid with 1-node halo
  DevParams dp;
  int ty, tz;        // tile rows, z-chunk length
  // 16-bit storage: the step's stochastic-rounding seed is seed (+ *seedp
  // when seedp is set: a device counter that captured graphs advance)
  uint32_t seed;
  const uint32_t* seedp;
  // open faces (inline moments stage only): 10-field set that receives the
  // post-moments state of the nodes next to a non-periodic wall, read by
  // the open-face correction after the launch; null without open faces
  void* sopen;
};

// Stage-path buffers (the reference's per-kernel arguments), as field sets.
struct StageArgs {
  void* S;           // 10-field set
  void* O;           // 10-field set
  const void* uin;   // 3-field set
  void* uout;        // 3-field set
  void* phiS;
  void* phiO;
  void* aux;         // 7-field set: n_x n_y n_z |g| theta_x theta_y theta_z
  void* lap;
  void* dvt;
  void* frc;         // 3-field set
  void* pre;         // 5-field set of the fast path (phase-field prep output)
  int gf_from_pre;   // PHI_int_boundary: donor gradient from pre instead of aux
  int div_only;      // phase prep: store div theta only (fused_grad_inline)
  void* rep;         // REPULSIVE_FLUX: 6-field set, locaux 3..8 (partner index, flux)
  int8_t* rep_mask;  // REPULSIVE_FLUX: nx*ny*nz
  const int8_t* isfh;
  DevParams dp;
  // open boundaries (openbc = 1): faces x(1) x(2) y(1) y(2) z(1) z(2), each
  // with its type (0 velocity, 1 pressure, 2 the node's own values) and its
  // pressure and velocity
  int obc;
  int obc_type[6];
  double obc_val[6][4];
};

// The fused kernel forms grad phi and lap phi itself from phi (the prep
// then stores only div theta) for the prefetching precisions without
// internal obstacles: with obstacles, PHI_int_boundary rewrites phi on solid
// nodes between the prep and the fused kernel, and the reference's gradients
// are those of the phi before it.
inline bool fused_grad_inline(int prec, const DevParams& dp) {
  return (prec == PREC_FP32 || prec == PREC_FP16S || prec == PREC_BF16S) && dp.twocomp && !dp.obst;
}

// Batched launches: one grid over many domains (same lattice, precision,
// component count and row tile; any sizes and physical parameters).
struct FusedBatchEntry {
  FusedArgs a;
  int tile0;  // first global tile of this domain
};
struct PrepBatchEntry {
  StageArgs a;
  int tz;     // z-chunk of the phase-field prep
  int tile0;
};

hipError_t launch_fused(const FusedArgs& a, int prec, bool inline_moments, hipStream_t s);
// every 16-bit scale equals def_scale(i): the kernels with compiled-in scales apply
bool default_scales(const DevParams& dp);
int fused_tile_count(const FusedArgs& a);
// prec may carry PREC_DEFSCALE (every domain with the default 16-bit scales)
hipError_t launch_fused_batch(const FusedBatchEntry* dev_tab, int ndom, int total_tiles, int prec, int lattice,
                              bool highorder, bool two, int ty, hipStream_t s);
int prep_tile_count(const StageArgs& a, int prec, int tz);
hipError_t launch_phase_prep_batch(const PrepBatchEntry* dev_tab, int ndom, int total_tiles, int prec, hipStream_t s);
hipError_t launch_stage(int stage, const StageArgs& a, int prec, hipStream_t s);

// storage conversion: host-order double/float array -> field (codec-aware)
hipError_t launch_phase_prep(const StageArgs& a, int prec, int tz, hipStream_t s);
hipError_t launch_encode_moments(void* S, const float* raw, const DevParams& dp, int prec, hipStream_t s);
hipError_t launch_decode_field(const void* S, int comp, void* out, int out_dtype, const DevParams& dp, int prec,
                               hipStream_t s);
// A field's place in a buffer: field f of an F-field set (F > 0), or the
// whole buffer as one x-fastest array (F == 0).
struct FieldLay {
  int F, f;
};
// host-order array (x fastest, nx*ny*nz) <-> stored field
hipError_t launch_convert_in(const void* src, int src_dtype, void* dst, int dst_kind, FieldLay dl, const DevParams& dp,
                             float scale, hipStream_t s);
hipError_t launch_convert_out(const void* src, int src_kind, FieldLay sl, void* dst, int dst_dtype, const DevParams& dp,
                              float invscale, hipStream_t s);
// nf consecutive fields between sets (element size esz bytes), node by node
hipError_t launch_set_copy(void* dst, FieldLay dl, const void* src, FieldLay sl, int nf, int esz, const DevParams& dp,
                           hipStream_t s);
hipError_t launch_reduce(const void* S, const void* phi, const int8_t* isfh, const DevParams& dp, int prec,
                         double* partial, int nblocks, double* out8, hipStream_t s);
// Reductions into the library's own device array (no allocation): at most
// REDUCE_SCRATCH_BLOCKS partial records of 8 doubles, then the 8 sums at
// double offset REDUCE_SCRATCH_SUMS.  reduce_scratch_ptr gives the array's
// device address on a device (looked up once per device).
constexpr int REDUCE_SCRATCH_BLOCKS = 1024;
constexpr size_t REDUCE_SCRATCH_SUMS = (size_t)REDUCE_SCRATCH_BLOCKS * 8;
constexpr int REDUCE_SCRATCH_MAX_DEVICES = 64;
hipError_t reduce_scratch_ptr(int device, double** out);
hipError_t launch_reduce_scratch(const void* S, const void* phi, const int8_t* isfh, const DevParams& dp, int prec,
                                 double* scratch_dev, hipStream_t s);
// BCPHIFLUX phase-mass correction: init = sum only (into lag[0]); otherwise
// sum, count, correction and its application to phi S
hipError_t launch_phi_lagrange(const StageArgs& a, int prec, double* partial, int nblocks, double* lag, bool init,
                               hipStream_t s);
hipError_t launch_fill_zero(void* p, size_t bytes, hipStream_t s);
// ctr[d] += 1 for d < n (the batch's per-domain step counters)
hipError_t launch_tick(uint32_t* ctr, int n, hipStream_t s);
hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s);

// element kinds for convert: 0 double, 1 float, 2 half (scaled), 3 bf16 (scaled)
enum Kind { K_F64 = 0, K_F32 = 1, K_F16 = 2, K_BF16 = 3 };

}  // namespace lbf
