// lbf_api.cpp -- the C-ABI of include/lbfast_hip.h.
//
// A context owns the state of one domain in HBM, laid out as structure of
// arrays: ten moment fields per ping-pong buffer (P, u, v, w, Pxx, Pyy,
// Pzz, Pxy, Pxz, Pyz; x fastest), the corrected-velocity ping-pong used as
// u_old, the phase field ping-pong, the two-component intermediates, and
// isfluid with a 1-node halo.  All launches go to one HIP stream.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstddef>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <string>
#include <vector>

#include "../../include/lbfast_hip.h"
#include "lbf_kernels.h"

using namespace lbf;

struct lbf_ctx {
  lbf_params p{};
  DevParams dp{};
  int prec = 0;
  hipStream_t stream = nullptr;
  int device = 0;
  size_t N = 0;
  size_t NS = 0;   // values per field of a field set: sx*ny*nz (row-interleaved sets)
  size_t tsz = 0;  // bytes of one moment value
  size_t asz = 0;  // bytes of one aux / phi / force value
  void* h[2] = {};    // 10-field sets (S/O ping-pong)
  void* U[2] = {};    // 3-field sets: corrected velocity (u_old ping-pong)
  void* phi[2] = {};
  void* aux = nullptr;  // 7-field set (stage path)
  void* lap = nullptr;
  void* dvt = nullptr;
  void* frc = nullptr;  // 3-field set (stage path)
  void* pre = nullptr;  // 5-field set (fast path phase-field prep)
  void* sopen = nullptr;  // open faces: post-moments state of the fast path (10-field set)
  int8_t* isfh = nullptr;
  void* rep = nullptr;         // REPULSIVE_FLUX: locaux 3..8 (6-field set)
  int8_t* rep_mask = nullptr;  // REPULSIVE_FLUX
  float* raw = nullptr;  // 10-field FP32 staging of 16-bit storage uploads
  bool raw_have[10] = {};
  void* staging = nullptr;
  double* partial = nullptr;
  double* red = nullptr;
  double* lag = nullptr;  // BCPHIFLUX: initial phi sum, last sum, count, correction
  int nred = 0;
  int cur = 0, ucur = 0;
  int step = 0;
  bool initialized = false;
  bool isf_set = false;
  bool raw_pending = false;
  int ty = 8, tz = 0, tz_prep = 64;
  std::vector<hipEvent_t> evs;  // per-launch event pairs (lbf_set_timing)
  int nev_used = 0;
  bool time_launches = false;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  bool timing_valid = false;
  std::vector<void*> allocs;
  std::string err;
};

static thread_local std::string g_err;

static int fail(lbf_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf; else g_err = buf;
  return code;
}
#define HIPCHK(c, expr)                                                                      \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return fail((c), LBF_EHIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

static void* dalloc(lbf_ctx* c, size_t bytes, hipError_t* e) {
  void* p = nullptr;
  *e = hipMalloc(&p, bytes);
  if (*e == hipSuccess) c->allocs.push_back(p);
  return p;
}

extern "C" {

// the layout the Python and Fortran mirrors of lbf_params assume
static_assert(sizeof(lbf_params) == 544, "lbf_params: 18 + 8 int32 and 33 + 24 double");
static_assert(offsetof(lbf_params, openbc) == 320 && offsetof(lbf_params, openbc_press) == 352,
              "lbf_params: open-boundary block");

int32_t lbf_abi_version(void) { return LBF_ABI_VERSION; }

void lbf_params_default(lbf_params* p) {
  if (!p) return;
  memset(p, 0, sizeof *p);
  p->nx = p->ny = p->nz = 64;
  p->lattice = 27;
  p->highorder = 1;
  p->pbc[0] = p->pbc[1] = p->pbc[2] = 1;
  p->precision = LBF_PREC_FP32;
  p->visc1 = p->visc2 = 0.1;
  p->rho_r = p->rho_b = 1.0;
  p->tau_diff = 0.1;
  p->width = 4.0;
  p->wettab_r = 90.0;
  lbf_params_derive(p);
}

// LBFAST.f90:212-219
int lbf_params_derive(lbf_params* p) {
  if (!p) return LBF_EINVAL;
  p->tau1 = p->visc1 * 3.0 + 0.5;
  p->omega = 1.0 / p->tau1;
  if (p->twocomponent) {
    if (p->width == 0.0) return LBF_EINVAL;
    p->beta = 12.0 * p->sigma / p->width;
    p->kapp = 1.5 * p->sigma * p->width;
    p->sharp_c = 4.0 * p->tau_diff / p->width;
    p->tau2 = p->visc2 * 3.0 + 0.5;
  }
  return LBF_OK;
}

// LBFAST.f90:258-328: walls at non-periodic faces (INTERNAL_OBSTACLES), halo
// copy (periodic image, else 1), then fluid nodes next to a non-fluid node
// on the compiled lattice become -1.
int lbf_classify_isfluid(const lbf_params* p, const int8_t* in, int8_t* out) {
  if (!p || !out) return LBF_EINVAL;
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  if (nx < 1 || ny < 1 || nz < 1) return LBF_EINVAL;
  auto at = [&](int i, int j, int k) { return (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k); };
  std::vector<int8_t> a((size_t)nx * ny * nz);
  for (size_t n = 0; n < a.size(); ++n) a[n] = in ? in[n] : 1;
  if (p->internal_obstacles) {
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          if (!p->pbc[0] && (i == 0 || i == nx - 1)) a[at(i, j, k)] = 0;
          if (!p->pbc[1] && (j == 0 || j == ny - 1)) a[at(i, j, k)] = 0;
          if (!p->pbc[2] && (k == 0 || k == nz - 1)) a[at(i, j, k)] = 0;
        }
  }
  // value seen through the halo: periodic image, or 1 past a non-periodic face
  auto look = [&](int i, int j, int k) -> int {
    const int n3[3] = {nx, ny, nz};
    int c[3] = {i, j, k};
    for (int d = 0; d < 3; ++d) {
      if (c[d] < 0 || c[d] >= n3[d]) {
        if (!p->pbc[d]) return 1;
        c[d] = (c[d] + n3[d]) % n3[d];
      }
    }
    return a[at(c[0], c[1], c[2])];
  };
  static const int EX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, -1, 1, 1, -1};
  static const int EY[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, -1, 1, -1, 1};
  static const int EZ[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1};
  const int nl = p->lattice == 27 ? 27 : 19;
  std::vector<int8_t> o(a);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        if (a[at(i, j, k)] != 1) continue;
        for (int l = 1; l < nl; ++l) {
          const int v = look(i + EX[l], j + EY[l], k + EZ[l]);
          if (v != 1 && v != -1) { o[at(i, j, k)] = -1; break; }
        }
      }
  memcpy(out, o.data(), o.size());
  return LBF_OK;
}

const char* lbf_global_last_error(void) { return g_err.c_str(); }
const char* lbf_last_error(const lbf_ctx* c) { return c ? c->err.c_str() : g_err.c_str(); }

static void fill_devparams(lbf_ctx* c) {
  const lbf_params& p = c->p;
  DevParams& d = c->dp;
  d.nx = p.nx; d.ny = p.ny; d.nz = p.nz;
  for (int i = 0; i < 3; ++i) d.pbc[i] = p.pbc[i];
  d.twocomp = p.twocomponent;
  d.dens = p.twocomponent && p.densratio;
  d.incomp = d.dens && p.interface_incomp;
  d.obst = p.internal_obstacles;
  d.buoy = p.buoyancy_forcing;
  d.lattice = p.lattice;
  d.highorder = p.highorder && p.lattice == 27;
  d.bb = p.internal_obstacles && p.bounce_back;
  d.smag = p.smagorinsky != 0;
  d.rep = p.twocomponent && p.repulsive_flux;
  d.nn = (size_t)p.nx * p.ny * p.nz;
  d.sx = set_sx(p.nx);
  // without DENSRATIO the reference's viscosity mix still reads rho_r/rho_b
  Phys<double>& q = d.pd;
  q.fx = p.fx; q.fy = p.fy; q.fz = p.fz;
  q.visc1 = p.visc1; q.visc2 = p.visc2; q.rho_r = p.rho_r; q.rho_b = p.rho_b;
  q.tau_diff = p.tau_diff; q.sharp_c = p.sharp_c; q.beta = p.beta; q.kapp = p.kapp;
  q.omega = p.omega; q.wettab = p.wettab_r;
  q.width = p.width; q.a_rep = p.a_rep; q.q_th = p.q_th; q.cos_opp = p.cos_opp_t;
  Phys<float>& f = d.pf;
  f.fx = (float)q.fx; f.fy = (float)q.fy; f.fz = (float)q.fz;
  f.visc1 = (float)q.visc1; f.visc2 = (float)q.visc2; f.rho_r = (float)q.rho_r; f.rho_b = (float)q.rho_b;
  f.tau_diff = (float)q.tau_diff; f.sharp_c = (float)q.sharp_c; f.beta = (float)q.beta; f.kapp = (float)q.kapp;
  f.omega = (float)q.omega; f.wettab = (float)q.wettab;
  f.width = (float)q.width; f.a_rep = (float)q.a_rep; f.q_th = (float)q.q_th; f.cos_opp = (float)q.cos_opp;
  // FP16S/BF16S: the given scales, or the power-of-two defaults (def_scale)
  for (int i = 0; i < 10; ++i) {
    const double s = p.store_scale[i] > 0 ? p.store_scale[i] : (double)def_scale(i);
    const bool devc = c->prec == LBF_PREC_FP16S || c->prec == LBF_PREC_BF16S;
    d.scale[i] = devc ? (float)s : 1.0f;
    d.invscale[i] = devc ? (float)(1.0 / s) : 1.0f;
  }
}

void lbf_destroy(lbf_ctx* c) {
  if (!c) return;
  if (c->stream) hipStreamSynchronize(c->stream);
  for (void* p : c->allocs) hipFree(p);
  for (hipEvent_t e : c->evs) hipEventDestroy(e);
  if (c->t0) hipEventDestroy(c->t0);
  if (c->t1) hipEventDestroy(c->t1);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

static void default_tiles(lbf_ctx* c) {
  // rows per block: 8 or 13, whichever needs fewer waves for the y extent
  // (a block of ty rows runs ty + 3 waves: two halo rows and the x-halo wave)
  const long w8 = (long)((c->p.ny + 7) / 8) * 11, w13 = (long)((c->p.ny + 12) / 13) * 16;
  c->ty = w13 < w8 ? 13 : 8;
  // z-chunks so the grid holds ~1024 blocks (4 per CU, one resident at a time)
  const long bxy = (long)((c->p.nx + 63) / 64) * ((c->p.ny + c->ty - 1) / c->ty);
  long nchunks = (1024 + bxy - 1) / bxy;
  if (nchunks < 1) nchunks = 1;
  long tz = (c->p.nz + nchunks - 1) / nchunks;
  if (tz < 8) tz = 8;
  if (tz > c->p.nz) tz = c->p.nz;
  c->tz = (int)tz;
  // phase prep: z-chunks so the grid holds at least ~2048 blocks. The 64 x 12
  // tiles of the 16/32-bit prep take chunks of 16..32 planes (measured at
  // 512^3: 1.27 ms with 32 planes against 1.38 ms with 128; each chunk re-reads
  // 4 planes of phi and recomputes 2 planes of theta); the FP64 prep's 64 x 8
  // tiles take chunks of at least 8
  StageArgs one{};
  one.dp = c->dp;
  const long pt = prep_tile_count(one, c->prec, c->p.nz);
  const bool tiles_z = pt > 0;
  const long pxy = tiles_z ? pt : (long)((c->p.nx + 63) / 64) * ((c->p.ny + 7) / 8);
  long pch = (2048 + pxy - 1) / pxy;
  if (pch < 1) pch = 1;
  long tzp = (c->p.nz + pch - 1) / pch;
  if (tiles_z && tzp > 32) tzp = 32;
  if (tzp < (tiles_z ? 16 : 8)) tzp = tiles_z ? 16 : 8;
  if (tzp > c->p.nz) tzp = c->p.nz;
  c->tz_prep = (int)tzp;
}

int lbf_create(const lbf_params* pin, lbf_ctx** out) {
  if (!pin || !out) return fail(nullptr, LBF_EINVAL, "null argument");
  *out = nullptr;
  lbf_params p = *pin;
  if (p.nx < 2 || p.ny < 2 || p.nz < 2) return fail(nullptr, LBF_ESIZE, "domain %dx%dx%d too small", p.nx, p.ny, p.nz);
  if ((double)p.nx * p.ny * p.nz > 4.0e9) return fail(nullptr, LBF_ESIZE, "domain too large");
  if (p.lattice != 19 && p.lattice != 27) return fail(nullptr, LBF_EINVAL, "LATTICE %d not supported (19, 27)", p.lattice);
  if (p.highorder && p.lattice != 27) return fail(nullptr, LBF_EINVAL, "HIGHORDER requires LATTICE 27");
  if (p.densratio && !p.twocomponent) return fail(nullptr, LBF_EINVAL, "DENSRATIO requires TWOCOMPONENT");
  if (p.precision < LBF_PREC_FP64 || p.precision > LBF_PREC_BF16S) return fail(nullptr, LBF_EINVAL, "bad precision");
  if (p.repulsive_flux && !p.twocomponent) return fail(nullptr, LBF_EINVAL, "REPULSIVE_FLUX requires TWOCOMPONENT");
  // bcs_mesoscopic_phifields, which holds the correction, is compiled only
  // with INTERNAL_OBSTACLES and TWOCOMPONENT (boundary_cds_module.f90:1345-1351)
  if (p.bcphiflux && !(p.twocomponent && p.internal_obstacles))
    return fail(nullptr, LBF_EINVAL, "BCPHIFLUX requires TWOCOMPONENT and INTERNAL_OBSTACLES");
  if (p.repulsive_flux && (p.precision == LBF_PREC_FP16S || p.precision == LBF_PREC_BF16S))
    return fail(nullptr, LBF_EINVAL, "REPULSIVE_FLUX runs the stage path: FP64, FP32 or MIXED storage");
  if (p.openbc) {
    // the open-boundary loops sit in bcs_mesoscopic_hfields / _phifields,
    // compiled only with INTERNAL_OBSTACLES (boundary_cds_module.f90:49-1311)
    if (!p.internal_obstacles) return fail(nullptr, LBF_EINVAL, "open boundaries require INTERNAL_OBSTACLES");
    if (p.precision == LBF_PREC_FP16S || p.precision == LBF_PREC_BF16S)
      return fail(nullptr, LBF_EINVAL, "open boundaries need FP64, FP32 or MIXED storage (no 16-bit codec for the open-face correction)");
    for (int f = 0; f < 6; ++f)
      if (p.openbc_type[f] < 0 || p.openbc_type[f] > 2)
        return fail(nullptr, LBF_EINVAL, "openbc_type %d of face %d: 0 (velocity), 1 (pressure) or 2 (local)",
                    p.openbc_type[f], f);
  }
  if (p.visc1 <= 0 || (p.twocomponent && p.visc2 <= 0)) return fail(nullptr, LBF_EINVAL, "viscosity must be > 0");
  if (p.twocomponent && p.densratio && (p.rho_r <= 0 || p.rho_b <= 0)) return fail(nullptr, LBF_EINVAL, "rho_r, rho_b must be > 0");
  if (lbf_params_derive(&p) != LBF_OK) return fail(nullptr, LBF_EINVAL, "width must be non-zero for TWOCOMPONENT");
  {
    // The kernels address one z plane of a field set with 32-bit byte offsets
    // and 32-bit buffer-resource ranges: the widest plane (the 10-field moment
    // set, or the 7-field aux set, or isfluid with its halo) must stay below 4 GiB.
    const uint64_t sx = (uint64_t)set_sx(p.nx), ny = (uint64_t)p.ny;
    const uint64_t tsz = p.precision == LBF_PREC_FP64 ? 8 : (p.precision == LBF_PREC_FP16S || p.precision == LBF_PREC_BF16S) ? 2 : 4;
    const uint64_t asz = p.precision == LBF_PREC_FP64 ? 8 : 4;
    uint64_t plane = 10 * sx * ny * tsz;
    if (7 * sx * ny * asz > plane) plane = 7 * sx * ny * asz;
    if ((uint64_t)(p.nx + 2) * (uint64_t)(p.ny + 2) > plane) plane = (uint64_t)(p.nx + 2) * (uint64_t)(p.ny + 2);
    if (plane + 4 > 0xFFFFFFFFull)
      return fail(nullptr, LBF_ESIZE,
                  "an x-y plane of %dx%d needs %llu bytes per field set, beyond the kernels' 32-bit plane offsets "
                  "(at most 4 GiB per plane; split the domain along x or y)",
                  p.nx, p.ny, (unsigned long long)plane);
  }

  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(nullptr, LBF_ENODEV, "no HIP device");
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return fail(nullptr, LBF_ENODEV, "hipGetDeviceProperties failed");
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(nullptr, LBF_ENODEV, "device %d is %s; this library is built for gfx950", dev, prop.gcnArchName);

  lbf_ctx* c = new lbf_ctx();
  c->p = p;
  c->prec = p.precision;
  c->device = dev;
  fill_devparams(c);
  c->N = (size_t)p.nx * p.ny * p.nz;
  c->NS = (size_t)set_sx(p.nx) * p.ny * p.nz;
  switch (c->prec) {
    case LBF_PREC_FP64: c->tsz = 8; c->asz = 8; break;
    case LBF_PREC_FP32: c->tsz = 4; c->asz = 4; break;
    case LBF_PREC_MIXED: c->tsz = 4; c->asz = 4; break;
    default: c->tsz = 2; c->asz = 4; break;
  }
  hipError_t e = hipSuccess;
  auto A = [&](size_t bytes) -> void* {
    if (e != hipSuccess) return nullptr;
    return dalloc(c, bytes, &e);
  };
  // field sets: F rows of sx values per (y, z) row (row-interleaved)
  for (int b = 0; b < 2; ++b) c->h[b] = A(10 * c->NS * c->tsz);
  for (int b = 0; b < 2; ++b) c->U[b] = A(3 * c->NS * c->tsz);
  if (p.twocomponent) {
    for (int b = 0; b < 2; ++b) c->phi[b] = A(c->N * c->asz);
    c->aux = A(7 * c->NS * c->asz);
    c->lap = A(c->N * c->asz);
    c->dvt = A(c->N * c->asz);
    c->pre = A(5 * c->NS * c->asz);
  }
  c->frc = A(3 * c->NS * c->asz);
  // open faces on the fast path: the fused kernel hands the open-face
  // correction the post-moments state of the wall-adjacent nodes here (a
  // whole set, so that it shares the moment sets' layout and offsets)
  if (p.openbc && p.internal_obstacles) c->sopen = A(10 * c->NS * c->tsz);
  if (c->dp.rep) {
    c->rep = A(6 * c->NS * c->asz);
    c->rep_mask = (int8_t*)A(c->N);
  }
  c->isfh = (int8_t*)A((size_t)(p.nx + 2) * (p.ny + 2) * (p.nz + 2));
  c->staging = A(c->N * 8);
  c->nred = REDUCE_SCRATCH_BLOCKS;  // the no-alloc variant's partials match this launch
  c->partial = (double*)A((size_t)c->nred * 8 * sizeof(double));
  c->red = (double*)A(8 * sizeof(double));
  if (p.bcphiflux) c->lag = (double*)A(4 * sizeof(double));
  if (e != hipSuccess) {
    std::string m = std::string("device allocation failed: ") + hipGetErrorString(e);
    lbf_destroy(c);
    return fail(nullptr, LBF_EHIP, "%s", m.c_str());
  }
  if (hipStreamCreate(&c->stream) != hipSuccess || hipEventCreate(&c->t0) != hipSuccess ||
      hipEventCreate(&c->t1) != hipSuccess) {
    lbf_destroy(c);
    return fail(nullptr, LBF_EHIP, "stream/event creation failed");
  }
  // all-fluid isfluid until the host uploads one (halo = 1 everywhere)
  std::vector<int8_t> ones((size_t)(p.nx + 2) * (p.ny + 2) * (p.nz + 2), 1);
  if (hipMemcpy(c->isfh, ones.data(), ones.size(), hipMemcpyHostToDevice) != hipSuccess) {
    lbf_destroy(c);
    return fail(nullptr, LBF_EHIP, "isfluid init failed");
  }
  default_tiles(c);
  *out = c;
  return LBF_OK;
}

int lbf_set_tile(lbf_ctx* c, int32_t ty, int32_t tz) {
  if (!c) return LBF_EINVAL;
  if (ty == 0 && tz == 0) { default_tiles(c); return LBF_OK; }
  if (ty != 0 && ty != 8 && ty != 13) return fail(c, LBF_EINVAL, "tile_y must be 8 or 13");
  if (ty) c->ty = ty;
  if (tz) {
    if (tz < 1) return fail(c, LBF_EINVAL, "tile_z must be >= 1");
    c->tz = tz > c->p.nz ? c->p.nz : tz;
  }
  return LBF_OK;
}

int32_t lbf_get_step(const lbf_ctx* c) { return c ? c->step : -1; }

// storage kind of a field id
static int field_kind(const lbf_ctx* c, int id, bool* moment) {
  *moment = (id >= 1 && id <= 10) || (id >= 27 && id <= 36);
  const bool devc = c->prec == LBF_PREC_FP16S || c->prec == LBF_PREC_BF16S;
  if (*moment || (id >= 12 && id <= 14)) {
    if (c->prec == LBF_PREC_FP64) return K_F64;
    if (c->prec == LBF_PREC_FP16S) return K_F16;
    if (c->prec == LBF_PREC_BF16S) return K_BF16;
    return K_F32;
  }
  (void)devc;
  return c->asz == 8 ? K_F64 : K_F32;
}

// Where a field id lives: the buffer and its place in it (field f of an
// F-field row-interleaved set, or F = 0 for a single x-fastest array).
struct FieldRef {
  void* base;
  FieldLay lay;
  size_t esz;  // bytes of one stored value
};
static FieldRef field_ref(lbf_ctx* c, int id) {
  const size_t T = c->tsz, A = c->asz;
  if (id >= 1 && id <= 10) return {c->h[c->cur], {10, id - 1}, T};
  if (id == LBF_FIELD_PHI) return {c->phi[c->cur], {0, 0}, A};
  if (id >= 12 && id <= 14) return {c->U[c->ucur], {3, id - 12}, T};
  if (id >= 15 && id <= 17) return {c->frc, {3, id - 15}, A};
  if (id >= 18 && id <= 24) return {c->aux, {7, id - 18}, A};
  if (id == LBF_FIELD_LAPPHI) return {c->lap, {0, 0}, A};
  if (id == LBF_FIELD_DIVTHN) return {c->dvt, {0, 0}, A};
  if (id >= 27 && id <= 36) return {c->h[1 - c->cur], {10, id - 27}, T};
  if (id == LBF_FIELD_PHI_O) return {c->phi[1 - c->cur], {0, 0}, A};
  if (id >= LBF_FIELD_LOCAUX3 && id <= LBF_FIELD_LOCAUX8) return {c->rep, {6, id - LBF_FIELD_LOCAUX3}, A};
  return {nullptr, {0, 0}, 0};
}

void* lbf_field_device_ptr(lbf_ctx* c, int id) {
  if (!c) return nullptr;
  const FieldRef r = field_ref(c, id);
  if (!r.base) return nullptr;
  // a set field: its first row; its rows recur every F*sx values
  return (char*)r.base + (size_t)r.lay.f * (size_t)c->dp.sx * r.esz;
}

int lbf_upload_field(lbf_ctx* c, int id, const void* host, int dtype) {
  if (!c || !host) return LBF_EINVAL;
  if (dtype != 0 && dtype != 1) return fail(c, LBF_EINVAL, "dtype must be 0 (float64) or 1 (float32)");
  const FieldRef dst = field_ref(c, id);
  if (!dst.base) return fail(c, LBF_EINVAL, "field %d not present in this configuration", id);
  bool mom;
  const int kind = field_kind(c, id, &mom);
  const bool devc = c->prec == LBF_PREC_FP16S || c->prec == LBF_PREC_BF16S;
  const size_t bytes = c->N * (dtype == 0 ? 8 : 4);
  HIPCHK(c, hipMemcpyAsync(c->staging, host, bytes, hipMemcpyHostToDevice, c->stream));
  if (mom && devc) {
    if (id > 10) return fail(c, LBF_EINVAL, "O-buffer upload not supported with 16-bit storage");
    // staged in FP32; encoded as deviations by lbf_init_state
    if (!c->raw) {
      hipError_t e;
      c->raw = (float*)dalloc(c, 10 * c->NS * 4, &e);  // a 10-field FP32 set
      if (e != hipSuccess) return fail(c, LBF_EHIP, "raw staging alloc failed");
    }
    HIPCHK(c, launch_convert_in(c->staging, dtype, c->raw, K_F32, FieldLay{10, id - 1}, c->dp, 1.0f, c->stream));
    c->raw_have[id - 1] = true;
    c->raw_pending = true;
  } else {
    float sc = 1.0f;
    if (devc && id >= 12 && id <= 14) sc = c->dp.scale[id - 11];
    HIPCHK(c, launch_convert_in(c->staging, dtype, dst.base, kind, dst.lay, c->dp, sc, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return LBF_OK;
}

int lbf_download_field(lbf_ctx* c, int id, void* host, int dtype) {
  if (!c || !host) return LBF_EINVAL;
  if (dtype != 0 && dtype != 1) return fail(c, LBF_EINVAL, "dtype must be 0 (float64) or 1 (float32)");
  const size_t bytes = c->N * (dtype == 0 ? 8 : 4);
  if (id == LBF_FIELD_RHOPHI) {
    // observable of copy_print (print_module.f90:734-760)
    std::vector<double> ph(c->N);
    if (c->p.twocomponent) {
      int rc = lbf_download_field(c, LBF_FIELD_PHI, ph.data(), 0);
      if (rc) return rc;
      for (size_t n = 0; n < c->N; ++n)
        if (c->p.densratio) ph[n] = c->p.rho_r * ph[n] + (1.0 - ph[n]) * c->p.rho_b;
    } else {
      int rc = lbf_download_field(c, LBF_FIELD_P, ph.data(), 0);
      if (rc) return rc;
    }
    if (dtype == 0) memcpy(host, ph.data(), bytes);
    else for (size_t n = 0; n < c->N; ++n) ((float*)host)[n] = (float)ph[n];
    return LBF_OK;
  }
  if (id == LBF_FIELD_REP_MASK) {
    if (!c->rep_mask) return fail(c, LBF_EINVAL, "field %d not present in this configuration", id);
    std::vector<int8_t> m(c->N);
    HIPCHK(c, hipMemcpyAsync(m.data(), c->rep_mask, c->N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t n = 0; n < c->N; ++n) {
      if (dtype == 0) ((double*)host)[n] = m[n];
      else ((float*)host)[n] = m[n];
    }
    return LBF_OK;
  }
  const FieldRef src = field_ref(c, id);
  if (!src.base) return fail(c, LBF_EINVAL, "field %d not present in this configuration", id);
  bool mom;
  const int kind = field_kind(c, id, &mom);
  const bool devc = c->prec == LBF_PREC_FP16S || c->prec == LBF_PREC_BF16S;
  if (mom) {
    if (devc && c->raw_pending) return fail(c, LBF_ESTATE, "moments uploaded but not encoded: call lbf_init_state");
    void* buf = (id <= 10) ? c->h[c->cur] : c->h[1 - c->cur];
    const int comp = (id <= 10) ? id - 1 : id - 27;
    HIPCHK(c, launch_decode_field(buf, comp, c->staging, dtype, c->dp, c->prec, c->stream));
  } else {
    float inv = 1.0f;
    if (devc && id >= 12 && id <= 14) inv = c->dp.invscale[id - 11];
    HIPCHK(c, launch_convert_out(src.base, kind, src.lay, c->staging, dtype, c->dp, inv, c->stream));
  }
  HIPCHK(c, hipMemcpyAsync(host, c->staging, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return LBF_OK;
}

int lbf_upload_isfluid(lbf_ctx* c, const int8_t* host) {
  if (!c || !host) return LBF_EINVAL;
  const int nx = c->p.nx, ny = c->p.ny, nz = c->p.nz;
  bool any_solid = false;
  for (size_t n = 0; n < c->N; ++n) {
    if (host[n] != 0 && host[n] != 1 && host[n] != -1) return fail(c, LBF_EINVAL, "isfluid values must be -1, 0 or 1");
    any_solid |= host[n] == 0;
  }
  if (any_solid && !c->p.internal_obstacles)
    return fail(c, LBF_EINVAL, "solid nodes require INTERNAL_OBSTACLES (internal_obstacles = 1)");
  const int hx = nx + 2, hy = ny + 2, hz = nz + 2;
  std::vector<int8_t> h((size_t)hx * hy * hz, 1);
  for (int k = 0; k < hz; ++k)
    for (int j = 0; j < hy; ++j)
      for (int i = 0; i < hx; ++i) {
        int c3[3] = {i - 1, j - 1, k - 1};
        const int n3[3] = {nx, ny, nz};
        bool one = false;
        for (int d = 0; d < 3; ++d) {
          if (c3[d] < 0 || c3[d] >= n3[d]) {
            if (!c->p.pbc[d]) one = true;
            c3[d] = (c3[d] + n3[d]) % n3[d];
          }
        }
        h[(size_t)i + (size_t)hx * ((size_t)j + (size_t)hy * k)] =
            one ? 1 : host[(size_t)c3[0] + (size_t)nx * ((size_t)c3[1] + (size_t)ny * c3[2])];
      }
  HIPCHK(c, hipMemcpy(c->isfh, h.data(), h.size(), hipMemcpyHostToDevice));
  c->isf_set = true;
  return LBF_OK;
}

static StageArgs stage_args(lbf_ctx* c) {
  StageArgs a{};
  a.S = c->h[c->cur];
  a.O = c->h[1 - c->cur];
  a.uin = c->U[c->ucur];
  a.uout = c->U[1 - c->ucur];
  a.phiS = c->phi[c->cur];
  a.phiO = c->phi[1 - c->cur];
  a.aux = c->aux;
  a.lap = c->lap;
  a.dvt = c->dvt;
  a.frc = c->frc;
  a.pre = c->pre;
  a.gf_from_pre = 0;
  a.div_only = fused_grad_inline(c->prec, c->dp) ? 1 : 0;
  a.rep = c->rep;
  a.rep_mask = c->rep_mask;
  a.isfh = c->isfh;
  a.dp = c->dp;
  a.obc = c->p.openbc != 0 && c->p.internal_obstacles != 0;
  for (int f = 0; f < 6; ++f) {
    a.obc_type[f] = c->p.openbc_type[f];
    a.obc_val[f][0] = c->p.openbc_press[f];
    a.obc_val[f][1] = c->p.openbc_u[f];
    a.obc_val[f][2] = c->p.openbc_v[f];
    a.obc_val[f][3] = c->p.openbc_w[f];
  }
  return a;
}

static FusedArgs fused_args(lbf_ctx* c) {
  FusedArgs a{};
  a.min = c->h[c->cur];
  a.mout = c->h[1 - c->cur];
  a.uin = c->U[c->ucur];
  a.uout = c->U[1 - c->ucur];
  a.phin = c->phi[c->cur];
  a.phiout = c->phi[1 - c->cur];
  a.pre = c->pre;
  a.frc = c->frc;
  a.isfh = c->isfh;
  a.dp = c->dp;
  a.ty = c->ty;
  a.tz = c->tz;
  a.seed = (uint32_t)c->step;
  a.seedp = nullptr;
  a.sopen = c->sopen;
  return a;
}

int lbf_init_state(lbf_ctx* c) {
  if (!c) return LBF_EINVAL;
  const bool devc = c->prec == LBF_PREC_FP16S || c->prec == LBF_PREC_BF16S;
  if (devc) {
    for (int f = 0; f < 10; ++f)
      if (!c->raw_have[f]) return fail(c, LBF_ESTATE, "moment field %d not uploaded", f + 1);
    HIPCHK(c, launch_encode_moments(c->h[c->cur], c->raw, c->dp, c->prec, c->stream));
    c->raw_pending = false;
  }
  // O := S (init_conditions_module.f90:584-587), u_old := u
  HIPCHK(c, launch_copy(c->h[1 - c->cur], c->h[c->cur], 10 * c->NS * c->tsz, c->stream));
  for (int b = 0; b < 2; ++b)
    HIPCHK(c, launch_set_copy(c->U[b], FieldLay{3, 0}, c->h[c->cur], FieldLay{10, 1}, 3, (int)c->tsz, c->dp, c->stream));
  if (c->p.twocomponent) {
    HIPCHK(c, launch_copy(c->phi[1 - c->cur], c->phi[c->cur], c->N * c->asz, c->stream));
    HIPCHK(c, launch_fill_zero(c->aux, 7 * c->NS * c->asz, c->stream));
    HIPCHK(c, launch_fill_zero(c->lap, c->N * c->asz, c->stream));
    HIPCHK(c, launch_fill_zero(c->dvt, c->N * c->asz, c->stream));
    HIPCHK(c, launch_fill_zero(c->pre, 5 * c->NS * c->asz, c->stream));
  }
  HIPCHK(c, launch_fill_zero(c->frc, 3 * c->NS * c->asz, c->stream));
  if (c->rep) {
    HIPCHK(c, launch_fill_zero(c->rep, 6 * c->NS * c->asz, c->stream));
    HIPCHK(c, launch_fill_zero(c->rep_mask, c->N, c->stream));
  }
  // bcs_phi before the loop (integrator_module.f90:179-182), with aux = 0;
  // under BCPHIFLUX the initial phase sum is taken first (initial conditions,
  // init_conditions_module.f90:469-490) and the correction follows the ghost update
  if (c->p.twocomponent && c->p.internal_obstacles) {
    StageArgs a = stage_args(c);
    if (c->lag) HIPCHK(c, launch_phi_lagrange(a, c->prec, c->partial, c->nred, c->lag, true, c->stream));
    HIPCHK(c, launch_stage(LBF_STAGE_PHI_INT_BOUNDARY, a, c->prec, c->stream));
    if (a.obc) HIPCHK(c, launch_stage(LBF_STAGE_OPEN_BC_PHI, a, c->prec, c->stream));
    if (c->lag) HIPCHK(c, launch_phi_lagrange(a, c->prec, c->partial, c->nred, c->lag, false, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->step = 0;
  c->initialized = true;
  return LBF_OK;
}

static int run_fused(lbf_ctx* c, bool inl) {
  FusedArgs a = fused_args(c);
  if (!c->time_launches) {
    HIPCHK(c, launch_fused(a, c->prec, inl, c->stream));
    return LBF_OK;
  }
  if (c->nev_used + 2 > (int)c->evs.size()) {
    for (int i = 0; i < 2; ++i) {
      hipEvent_t ev;
      HIPCHK(c, hipEventCreate(&ev));
      c->evs.push_back(ev);
    }
  }
  HIPCHK(c, hipEventRecord(c->evs[c->nev_used], c->stream));
  HIPCHK(c, launch_fused(a, c->prec, inl, c->stream));
  HIPCHK(c, hipEventRecord(c->evs[c->nev_used + 1], c->stream));
  c->nev_used += 2;
  return LBF_OK;
}

// One time step as the reference's stage sequence (integrator_module.f90:266-545),
// stream-ordered, no host synchronisation.
static int staged_step(lbf_ctx* c) {
  const bool two = c->p.twocomponent != 0, obst = c->p.internal_obstacles != 0;
  StageArgs a = stage_args(c);
  if (two) {
    HIPCHK(c, launch_stage(LBF_STAGE_NORM, a, c->prec, c->stream));
    if (obst) HIPCHK(c, launch_stage(LBF_STAGE_PHI_INT_BOUNDARY, a, c->prec, c->stream));
    if (a.obc) HIPCHK(c, launch_stage(LBF_STAGE_OPEN_BC_PHI, a, c->prec, c->stream));
    if (c->lag) HIPCHK(c, launch_phi_lagrange(a, c->prec, c->partial, c->nred, c->lag, false, c->stream));
    HIPCHK(c, launch_stage(LBF_STAGE_DIV_THETA_N, a, c->prec, c->stream));
    if (c->dp.rep) {
      HIPCHK(c, launch_stage(LBF_STAGE_THINFILM_SCAN, a, c->prec, c->stream));
      HIPCHK(c, launch_stage(LBF_STAGE_REPULSIVE_FLUX, a, c->prec, c->stream));
    }
  }
  HIPCHK(c, launch_stage(LBF_STAGE_MOMENTS, a, c->prec, c->stream));
  int rc = run_fused(c, false);
  if (rc) return rc;
  if (two) HIPCHK(c, launch_stage(LBF_STAGE_UPDATE_PHI, a, c->prec, c->stream));
  if (obst && !c->p.bounce_back) HIPCHK(c, launch_stage(LBF_STAGE_LB_INT_BOUNDARY, a, c->prec, c->stream));
  // the open faces follow LB_int_boundary in bcs_mesoscopic_hfields, also
  // under BOUNCE_BACK (boundary_cds_module.f90:49-57)
  if (a.obc) HIPCHK(c, launch_stage(LBF_STAGE_OPEN_BC_H, a, c->prec, c->stream));
  c->cur ^= 1;
  c->ucur ^= 1;
  return LBF_OK;
}

int lbf_step(lbf_ctx* c, int32_t nsteps) {
  if (!c) return LBF_EINVAL;
  if (!c->initialized) return fail(c, LBF_ESTATE, "lbf_init_state must be called before lbf_step");
  if (nsteps < 0) return fail(c, LBF_EINVAL, "nsteps < 0");
  c->nev_used = 0;
  HIPCHK(c, hipEventRecord(c->t0, c->stream));
  for (int s = 0; s < nsteps; ++s) {
    ++c->step;
    if (c->dp.rep) {
      // the repulsive scan reads the normals of compute_norm's aux set, which
      // the fused fast path does not form: run the reference's stage order
      int rc = staged_step(c);
      if (rc) return rc;
      continue;
    }
    const bool obc = c->sopen != nullptr;
    if (c->p.twocomponent) {
      // compute_norm + div_theta_n (one tiled pass), then the ghost-phi
      // update on the solid nodes (and the open faces' wall copies), then
      // the fused step
      StageArgs a = stage_args(c);
      a.gf_from_pre = 1;
      HIPCHK(c, launch_phase_prep(a, c->prec, c->tz_prep, c->stream));
      if (c->p.internal_obstacles) HIPCHK(c, launch_stage(LBF_STAGE_PHI_INT_BOUNDARY, a, c->prec, c->stream));
      if (obc) HIPCHK(c, launch_stage(LBF_STAGE_OPEN_BC_PHI, a, c->prec, c->stream));
      if (c->lag) HIPCHK(c, launch_phi_lagrange(a, c->prec, c->partial, c->nred, c->lag, false, c->stream));
    }
    int rc = run_fused(c, true);
    if (rc) return rc;
    if (obc) {
      // the open faces correct the fused output (which already holds the
      // near-wall closure) from the post-moments state the kernel set aside
      StageArgs a = stage_args(c);
      a.S = c->sopen;
      HIPCHK(c, launch_stage(LBF_STAGE_OPEN_BC_H, a, c->prec, c->stream));
    }
    c->cur ^= 1;
    c->ucur ^= 1;
  }
  HIPCHK(c, hipEventRecord(c->t1, c->stream));
  c->timing_valid = nsteps > 0;
  return LBF_OK;
}

int lbf_last_step_timing(lbf_ctx* c, double* total_ms, double* fused_ms) {
  if (!c || !total_ms || !fused_ms) return LBF_EINVAL;
  if (!c->timing_valid) return fail(c, LBF_ESTATE, "no timed lbf_step call");
  HIPCHK(c, hipEventSynchronize(c->t1));
  float ms = 0;
  HIPCHK(c, hipEventElapsedTime(&ms, c->t0, c->t1));
  *total_ms = ms;
  if (!c->time_launches) {
    *fused_ms = -1.0;  // per-launch timing is off
    return LBF_OK;
  }
  double f = 0;
  for (int i = 0; i + 1 < c->nev_used; i += 2) {
    float m2 = 0;
    HIPCHK(c, hipEventElapsedTime(&m2, c->evs[i], c->evs[i + 1]));
    f += m2;
  }
  *fused_ms = f;
  return LBF_OK;
}

int lbf_set_timing(lbf_ctx* c, int32_t on) {
  if (!c) return LBF_EINVAL;
  c->time_launches = on != 0;
  c->nev_used = 0;
  c->timing_valid = false;
  return LBF_OK;
}

int lbf_begin_step(lbf_ctx* c) {
  if (!c) return LBF_EINVAL;
  if (!c->initialized) return fail(c, LBF_ESTATE, "lbf_init_state must be called first");
  ++c->step;
  return LBF_OK;
}

int lbf_stage(lbf_ctx* c, int id) {
  if (!c) return LBF_EINVAL;
  if (!c->initialized) return fail(c, LBF_ESTATE, "lbf_init_state must be called first");
  const bool two = c->p.twocomponent != 0, obst = c->p.internal_obstacles != 0;
  switch (id) {
    case LBF_STAGE_NORM:
    case LBF_STAGE_DIV_THETA_N:
    case LBF_STAGE_UPDATE_PHI:
      if (!two) return LBF_OK;
      break;
    case LBF_STAGE_PHI_INT_BOUNDARY:
      if (!two || !obst) return LBF_OK;
      break;
    case LBF_STAGE_LB_INT_BOUNDARY:
      if (!obst || c->p.bounce_back) return LBF_OK;  // boundary_cds_module.f90:49-56
      break;
    case LBF_STAGE_MOMENTS:
      break;
    case LBF_STAGE_THINFILM_SCAN:
    case LBF_STAGE_REPULSIVE_FLUX:
      if (!c->dp.rep) return LBF_OK;
      break;
    case LBF_STAGE_OPEN_BC_PHI:
      if (!two || !obst || !c->p.openbc) return LBF_OK;
      break;
    case LBF_STAGE_OPEN_BC_H:
      if (!obst || !c->p.openbc) return LBF_OK;
      break;
    case LBF_STAGE_PHI_LAGRANGE: {
      if (!c->lag) return LBF_OK;
      StageArgs a = stage_args(c);
      HIPCHK(c, launch_phi_lagrange(a, c->prec, c->partial, c->nred, c->lag, false, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      return LBF_OK;
    }
    case LBF_STAGE_FUSED: {
      if (c->prec == LBF_PREC_FP16S || c->prec == LBF_PREC_BF16S)
        return fail(c, LBF_EINVAL, "the stage-path fused kernel supports FP64/FP32/MIXED storage");
      c->nev_used = 0;
      int rc = run_fused(c, false);
      if (rc) return rc;
      HIPCHK(c, hipStreamSynchronize(c->stream));
      return LBF_OK;
    }
    default:
      return fail(c, LBF_EINVAL, "unknown stage %d", id);
  }
  StageArgs a = stage_args(c);
  HIPCHK(c, launch_stage(id, a, c->prec, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return LBF_OK;
}

int lbf_end_step(lbf_ctx* c) {
  if (!c) return LBF_EINVAL;
  c->cur ^= 1;
  c->ucur ^= 1;
  return LBF_OK;
}

int lbf_reduce_fp64(lbf_ctx* c, double out[8]) {
  if (!c || !out) return LBF_EINVAL;
  HIPCHK(c, launch_reduce(c->h[c->cur], c->phi[c->cur], c->isfh, c->dp, c->prec, c->partial, c->nred, c->red, c->stream));
  HIPCHK(c, hipMemcpyAsync(out, c->red, 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return LBF_OK;
}

int lbf_reduce_fp64_noalloc(lbf_ctx* c, double out[8]) {
  if (!c || !out) return LBF_EINVAL;
  double* scr = nullptr;
  HIPCHK(c, reduce_scratch_ptr(c->device, &scr));
  HIPCHK(c, launch_reduce_scratch(c->h[c->cur], c->phi[c->cur], c->isfh, c->dp, c->prec, scr, c->stream));
  HIPCHK(c, hipMemcpyAsync(out, scr + REDUCE_SCRATCH_SUMS, 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return LBF_OK;
}

int lbf_synchronize(lbf_ctx* c) {
  if (!c) return LBF_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return LBF_OK;
}

// ---------------------------------------------------------------------------
// Batches: many independent domains advanced by one prep launch and one fused
// launch per step (or a captured graph of such steps).  The contexts keep
// their state and stay usable for uploads/downloads between batch calls.
struct lbf_batch {
  std::vector<lbf_ctx*> ctx;
  int prec = 0, lattice = 27;
  bool ho = false, two = false;
  int ty = 13;
  std::vector<int> tz, tz_prep;
  // table position -> domain: the order in which the domains' tiles fill the
  // grid (each XCD takes a contiguous eighth of it)
  std::vector<int> order;
  FusedBatchEntry* ftab[2] = {};
  PrepBatchEntry* ptab[2] = {};
  int ftotal[2] = {}, ptotal[2] = {};
  int par = 0;  // 0: tables built from the contexts' current buffers
  std::vector<int> tcur, tucur, tstep;  // each context's (cur, ucur, step) that table 0 describes
  uint32_t* ctr = nullptr;  // per-domain step counters: the 16-bit codecs' rounding seeds
  hipStream_t stream = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  std::vector<hipEvent_t> cev;  // one per context: orders its stream before batch work
  bool timing_valid = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_steps = 0;
  std::string err;
};

static int bfail(lbf_batch* b, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (b) b->err = buf; else g_err = buf;
  return code;
}
#define BHIPCHK(b, expr)                                                                     \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return bfail((b), LBF_EHIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

static void batch_drop_graph(lbf_batch* b) {
  if (b->gexec) hipGraphExecDestroy(b->gexec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  b->graph_steps = 0;
}

// Device tables for both buffer parities (parity 1 = after one step).
static int batch_tables(lbf_batch* b) {
  const int nd = (int)b->ctx.size();
  for (int par = 0; par < 2; ++par) {
    std::vector<FusedBatchEntry> f(nd);
    std::vector<PrepBatchEntry> q(nd);
    int ft = 0, pt = 0;
    for (int i = 0; i < nd; ++i) {
      const int d = b->order[i];
      lbf_ctx* c = b->ctx[d];
      const int cur0 = c->cur, ucur0 = c->ucur;
      c->cur ^= par;
      c->ucur ^= par;
      f[i].a = fused_args(c);
      f[i].a.ty = b->ty;
      f[i].a.tz = b->tz[d];
      // the seed of a step is the domain's step number, as in lbf_step:
      // 1 + the counter, which the step's tick advances afterwards
      f[i].a.seed = 1u;
      f[i].a.seedp = b->ctr + d;
      f[i].tile0 = ft;
      ft += fused_tile_count(f[i].a);
      q[i].a = stage_args(c);
      q[i].a.gf_from_pre = 1;
      q[i].tz = b->tz_prep[d];
      q[i].tile0 = pt;
      if (b->two) pt += prep_tile_count(q[i].a, b->prec, b->tz_prep[d]);
      c->cur = cur0;
      c->ucur = ucur0;
    }
    if (!b->ftab[par]) BHIPCHK(b, hipMalloc(&b->ftab[par], nd * sizeof(FusedBatchEntry)));
    if (!b->ptab[par]) BHIPCHK(b, hipMalloc(&b->ptab[par], nd * sizeof(PrepBatchEntry)));
    BHIPCHK(b, hipMemcpy(b->ftab[par], f.data(), nd * sizeof(FusedBatchEntry), hipMemcpyHostToDevice));
    BHIPCHK(b, hipMemcpy(b->ptab[par], q.data(), nd * sizeof(PrepBatchEntry), hipMemcpyHostToDevice));
    b->ftotal[par] = ft;
    b->ptotal[par] = pt;
  }
  b->tcur.resize(nd);
  b->tucur.resize(nd);
  b->tstep.resize(nd);
  std::vector<uint32_t> steps(nd);
  for (int d = 0; d < nd; ++d) {
    b->tcur[d] = b->ctx[d]->cur;
    b->tucur[d] = b->ctx[d]->ucur;
    b->tstep[d] = b->ctx[d]->step;
    steps[d] = (uint32_t)b->ctx[d]->step;
  }
  BHIPCHK(b, hipMemcpy(b->ctr, steps.data(), nd * sizeof(uint32_t), hipMemcpyHostToDevice));
  b->par = 0;
  return LBF_OK;
}

// A member context stepped on its own (lbf_step, the stage path) since the
// tables were built has swapped its buffers: rebuild the tables from the
// contexts' current state and drop a captured graph that refers to the old ones.
static int batch_refresh(lbf_batch* b) {
  bool stale = false;
  for (size_t d = 0; d < b->ctx.size(); ++d)
    stale |= b->ctx[d]->cur != b->tcur[d] || b->ctx[d]->ucur != b->tucur[d] || b->ctx[d]->step != b->tstep[d];
  if (!stale) return LBF_OK;
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  batch_drop_graph(b);
  return batch_tables(b);
}

// Enqueue one step on b->stream at the batch's current parity (no state change).
static int batch_enqueue(lbf_batch* b, int par) {
  const int nd = (int)b->ctx.size();
  if (b->two) BHIPCHK(b, launch_phase_prep_batch(b->ptab[par], nd, b->ptotal[par], b->prec, b->stream));
  int kprec = b->prec;
  if (kprec == LBF_PREC_FP16S || kprec == LBF_PREC_BF16S) {
    bool ds = true;
    for (lbf_ctx* c : b->ctx) ds = ds && default_scales(c->dp);
    if (ds) kprec |= PREC_DEFSCALE;
  }
  BHIPCHK(b, launch_fused_batch(b->ftab[par], nd, b->ftotal[par], kprec, b->lattice, b->ho, b->two, b->ty, b->stream));
  if (b->prec == LBF_PREC_FP16S || b->prec == LBF_PREC_BF16S) BHIPCHK(b, launch_tick(b->ctr, nd, b->stream));
  return LBF_OK;
}

static void batch_advance(lbf_batch* b, int nsteps) {
  for (lbf_ctx* c : b->ctx) {
    if (nsteps & 1) {
      c->cur ^= 1;
      c->ucur ^= 1;
    }
    c->step += nsteps;
  }
}

int lbf_batch_create(lbf_ctx* const* ctxs, int32_t ndomains, lbf_batch** out) {
  if (!ctxs || !out || ndomains < 1) return bfail(nullptr, LBF_EINVAL, "bad arguments");
  *out = nullptr;
  lbf_ctx* c0 = ctxs[0];
  if (!c0) return bfail(nullptr, LBF_EINVAL, "null context");
  for (int d = 0; d < ndomains; ++d) {
    lbf_ctx* c = ctxs[d];
    if (!c || !c->initialized) return bfail(nullptr, LBF_ESTATE, "domain %d: context missing or not initialised", d);
    if (c->prec != c0->prec || c->p.lattice != c0->p.lattice || c->dp.highorder != c0->dp.highorder ||
        c->p.twocomponent != c0->p.twocomponent || c->device != c0->device)
      return bfail(nullptr, LBF_EINVAL, "domain %d: a batch shares lattice, HIGHORDER, precision, component count and device", d);
    if (c->p.internal_obstacles || c->p.repulsive_flux)
      return bfail(nullptr, LBF_EINVAL, "domain %d: obstacles and REPULSIVE_FLUX run per domain, not batched", d);
    if (c->p.twocomponent && c->prec == LBF_PREC_FP64)
      return bfail(nullptr, LBF_EINVAL, "two-component batches need float phase-field storage (not FP64)");
    for (int e = 0; e < d; ++e)
      if (ctxs[e] == c) return bfail(nullptr, LBF_EINVAL, "domain %d repeats domain %d", d, e);
  }
  lbf_batch* b = new lbf_batch();
  b->ctx.assign(ctxs, ctxs + ndomains);
  b->prec = c0->prec;
  b->lattice = c0->p.lattice;
  b->ho = c0->dp.highorder != 0;
  b->two = c0->p.twocomponent != 0;
  b->ty = 13;
  for (lbf_ctx* c : b->ctx) {
    b->tz.push_back(c->p.nz < 32 ? c->p.nz : 32);
    b->tz_prep.push_back(c->p.nz < 32 ? c->p.nz : 32);
  }
  for (int d = 0; d < ndomains; ++d) b->order.push_back(d);
  bool ok = hipStreamCreate(&b->stream) == hipSuccess && hipEventCreate(&b->t0) == hipSuccess &&
            hipEventCreate(&b->t1) == hipSuccess && hipMalloc(&b->ctr, ndomains * sizeof(uint32_t)) == hipSuccess;
  for (lbf_ctx* c : b->ctx) {
    hipEvent_t e = nullptr;
    ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
    b->cev.push_back(e);
    ok = ok && hipStreamSynchronize(c->stream) == hipSuccess;
  }
  if (!ok) {
    lbf_batch_destroy(b);
    return bfail(nullptr, LBF_EHIP, "stream/event creation failed");
  }
  int rc = batch_tables(b);
  if (rc) {
    g_err = b->err;
    lbf_batch_destroy(b);
    return rc;
  }
  *out = b;
  return LBF_OK;
}

void lbf_batch_destroy(lbf_batch* b) {
  if (!b) return;
  if (b->stream) hipStreamSynchronize(b->stream);
  batch_drop_graph(b);
  for (int i = 0; i < 2; ++i) {
    if (b->ftab[i]) hipFree(b->ftab[i]);
    if (b->ptab[i]) hipFree(b->ptab[i]);
  }
  if (b->ctr) hipFree(b->ctr);
  if (b->t0) hipEventDestroy(b->t0);
  if (b->t1) hipEventDestroy(b->t1);
  for (hipEvent_t e : b->cev)
    if (e) hipEventDestroy(e);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
}

const char* lbf_batch_last_error(const lbf_batch* b) { return b ? b->err.c_str() : g_err.c_str(); }

int lbf_batch_set_tile(lbf_batch* b, int32_t ty, int32_t tz, int32_t tz_prep) {
  if (!b) return LBF_EINVAL;
  if (ty != 8 && ty != 13) return bfail(b, LBF_EINVAL, "tile_y must be 8 or 13");
  if (tz < 1 || tz_prep < 1) return bfail(b, LBF_EINVAL, "z chunks must be >= 1");
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  b->ty = ty;
  for (size_t d = 0; d < b->ctx.size(); ++d) {
    const int nz = b->ctx[d]->p.nz;
    b->tz[d] = tz < nz ? tz : nz;
    b->tz_prep[d] = tz_prep < nz ? tz_prep : nz;
  }
  batch_drop_graph(b);
  // tables describe the buffers at the batch's parity 0: rebuild from the contexts' state
  b->par = 0;
  return batch_tables(b);
}

int lbf_batch_step(lbf_batch* b, int32_t nsteps) {
  if (!b) return LBF_EINVAL;
  if (nsteps < 0) return bfail(b, LBF_EINVAL, "nsteps < 0");
  {
    int rc = batch_refresh(b);
    if (rc) return rc;
  }
  // work queued on a domain's own stream (uploads) comes first ...
  for (size_t d = 0; d < b->ctx.size(); ++d) {
    BHIPCHK(b, hipEventRecord(b->cev[d], b->ctx[d]->stream));
    BHIPCHK(b, hipStreamWaitEvent(b->stream, b->cev[d], 0));
  }
  BHIPCHK(b, hipEventRecord(b->t0, b->stream));
  int done = 0;
  if (b->gexec && b->graph_steps > 0) {
    while (nsteps - done >= b->graph_steps) {
      BHIPCHK(b, hipGraphLaunch(b->gexec, b->stream));
      done += b->graph_steps;  // even: parity unchanged
    }
  }
  for (; done < nsteps; ++done) {
    int rc = batch_enqueue(b, b->par);
    if (rc) return rc;
    b->par ^= 1;
  }
  BHIPCHK(b, hipEventRecord(b->t1, b->stream));
  // ... and work later queued on it (downloads, single-domain steps) comes after
  for (lbf_ctx* c : b->ctx) BHIPCHK(b, hipStreamWaitEvent(c->stream, b->t1, 0));
  batch_advance(b, nsteps);
  b->timing_valid = nsteps > 0;
  // keep tables indexed from the contexts' current buffers
  if (b->par) {
    FusedBatchEntry* f = b->ftab[0]; b->ftab[0] = b->ftab[1]; b->ftab[1] = f;
    PrepBatchEntry* q = b->ptab[0]; b->ptab[0] = b->ptab[1]; b->ptab[1] = q;
    int t = b->ftotal[0]; b->ftotal[0] = b->ftotal[1]; b->ftotal[1] = t;
    t = b->ptotal[0]; b->ptotal[0] = b->ptotal[1]; b->ptotal[1] = t;
    b->par = 0;
    batch_drop_graph(b);  // a captured graph refers to the old parity's tables
  }
  for (size_t d = 0; d < b->ctx.size(); ++d) {
    b->tcur[d] = b->ctx[d]->cur;
    b->tucur[d] = b->ctx[d]->ucur;
    b->tstep[d] = b->ctx[d]->step;
  }
  return LBF_OK;
}

int lbf_batch_capture(lbf_batch* b, int32_t nsteps) {
  if (!b) return LBF_EINVAL;
  if (nsteps < 2 || (nsteps & 1)) return bfail(b, LBF_EINVAL, "a captured graph holds an even number of steps (>= 2)");
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  {
    int rc = batch_refresh(b);
    if (rc) return rc;
  }
  batch_drop_graph(b);
  BHIPCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int s = 0; s < nsteps && !rc; ++s) rc = batch_enqueue(b, s & 1);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return bfail(b, LBF_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  b->graph = g;
  e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    batch_drop_graph(b);
    return bfail(b, LBF_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  }
  b->graph_steps = nsteps;
  return LBF_OK;
}

int lbf_batch_synchronize(lbf_batch* b) {
  if (!b) return LBF_EINVAL;
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  return LBF_OK;
}

int lbf_batch_timing(lbf_batch* b, double* ms) {
  if (!b || !ms) return LBF_EINVAL;
  if (!b->timing_valid) return bfail(b, LBF_ESTATE, "no timed lbf_batch_step call");
  BHIPCHK(b, hipEventSynchronize(b->t1));
  float m = 0;
  BHIPCHK(b, hipEventElapsedTime(&m, b->t0, b->t1));
  *ms = m;
  return LBF_OK;
}

// The project's autotuner: times `trial_steps` (even) steps of each candidate
// on the batch itself, keeps the fastest, and restores every domain's state
// afterwards.  Candidates, searched in turn: the row tile (8 or 13) with one z
// chunk (16..128) for every domain; then the z chunk of each size class
// (domains of one nz) with the others held; then the table order (as given,
// largest first, or largest and smallest alternating).
int lbf_batch_autotune(lbf_batch* b, int32_t trial_steps, int32_t* best_ty, int32_t* best_tz) {
  if (!b) return LBF_EINVAL;
  if (trial_steps < 2 || (trial_steps & 1)) return bfail(b, LBF_EINVAL, "trial_steps must be even and >= 2");
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  {
    int rc = batch_refresh(b);
    if (rc) return rc;
  }
  for (lbf_ctx* c : b->ctx) BHIPCHK(b, hipStreamSynchronize(c->stream));
  // snapshot: the S moments, u_old and phi each step reads
  std::vector<void*> snap;
  auto release = [&]() { for (void* q : snap) hipFree(q); snap.clear(); };
  for (lbf_ctx* c : b->ctx) {
    const size_t bytes[3] = {10 * c->NS * c->tsz, 3 * c->NS * c->tsz, c->p.twocomponent ? c->N * c->asz : 0};
    void* src[3] = {c->h[c->cur], c->U[c->ucur], c->p.twocomponent ? c->phi[c->cur] : nullptr};
    for (int i = 0; i < 3; ++i) {
      void* q = nullptr;
      if (bytes[i]) {
        hipError_t e = hipMalloc(&q, bytes[i]);
        if (e != hipSuccess) { release(); return bfail(b, LBF_EHIP, "autotune snapshot: %s", hipGetErrorString(e)); }
        e = hipMemcpyAsync(q, src[i], bytes[i], hipMemcpyDeviceToDevice, b->stream);
        if (e != hipSuccess) { hipFree(q); release(); return bfail(b, LBF_EHIP, "autotune snapshot copy"); }
      }
      snap.push_back(q);
    }
  }
  // one timed trial of the batch's current configuration (warm-up pair, then
  // trial_steps steps; even counts keep the buffer parities)
  auto trial = [&](double* ms) -> int {
    int rc = batch_tables(b);
    for (int s = 0; s < 2 && !rc; ++s) rc = batch_enqueue(b, s & 1);
    if (rc) return rc;
    BHIPCHK(b, hipEventRecord(b->t0, b->stream));
    for (int s = 0; s < trial_steps && !rc; ++s) rc = batch_enqueue(b, s & 1);
    if (rc) return rc;
    BHIPCHK(b, hipEventRecord(b->t1, b->stream));
    BHIPCHK(b, hipEventSynchronize(b->t1));
    float m = 0;
    BHIPCHK(b, hipEventElapsedTime(&m, b->t0, b->t1));
    *ms = m;
    return LBF_OK;
  };
  const int nd = (int)b->ctx.size();
  auto set_uniform = [&](int tz) {
    for (int d = 0; d < nd; ++d) {
      const int nz = b->ctx[d]->p.nz;
      b->tz[d] = b->tz_prep[d] = tz < nz ? tz : nz;
    }
  };
  batch_drop_graph(b);
  // 1. row tile and one z chunk for every domain
  const int tys[2] = {8, 13}, tzs[4] = {16, 32, 64, 128};
  double best = 1e300;
  int bty = b->ty, btz = b->tz.empty() ? 32 : b->tz[0];
  for (int ty : tys)
    for (int tz : tzs) {
      b->ty = ty;
      set_uniform(tz);
      double m = 0;
      int rc = trial(&m);
      if (rc) { release(); return rc; }
      if (m < best) { best = m; bty = ty; btz = tz; }
    }
  b->ty = bty;
  set_uniform(btz);
  // 2. the z chunk of each size class (domains of one nz), the others held:
  // short domains lose least to the chunks' two halo planes with one chunk,
  // tall ones may balance better with more
  std::vector<int> classes;
  for (lbf_ctx* c : b->ctx)
    if (std::find(classes.begin(), classes.end(), c->p.nz) == classes.end()) classes.push_back(c->p.nz);
  if (classes.size() > 1) {
    for (int nz : classes) {
      const int cand[5] = {16, 32, 64, 128, nz};
      int keep = -1;
      for (int tz : cand) {
        if (tz > nz) continue;
        std::vector<int> save = b->tz;
        for (int d = 0; d < nd; ++d)
          if (b->ctx[d]->p.nz == nz) b->tz[d] = b->tz_prep[d] = tz;
        double m = 0;
        int rc = trial(&m);
        if (rc) { release(); return rc; }
        if (m < best * 0.995) { best = m; keep = tz; }
        b->tz = save;
        b->tz_prep = save;
      }
      if (keep > 0)
        for (int d = 0; d < nd; ++d)
          if (b->ctx[d]->p.nz == nz) b->tz[d] = b->tz_prep[d] = keep;
    }
  }
  // 3. table order: as given, largest first (the tallest work items start
  // first), or largest and smallest alternating (each XCD's contiguous eighth
  // of the grid holds a mix)
  {
    std::vector<int> by(nd);
    for (int d = 0; d < nd; ++d) by[d] = d;
    std::stable_sort(by.begin(), by.end(), [&](int x, int y) { return b->ctx[x]->N > b->ctx[y]->N; });
    std::vector<int> alt;
    for (int lo = 0, hi = nd - 1; lo <= hi;) {
      alt.push_back(by[lo++]);
      if (lo <= hi) alt.push_back(by[hi--]);
    }
    const std::vector<int> given = b->order;
    std::vector<int> bestord = given;
    for (const std::vector<int>* o : {&by, &alt}) {
      b->order = *o;
      double m = 0;
      int rc = trial(&m);
      if (rc) { b->order = given; release(); return rc; }
      if (m < best * 0.995) { best = m; bestord = *o; }
    }
    b->order = bestord;
  }
  // restore the snapshot (trial steps were even: the parities are unchanged)
  size_t k = 0;
  for (lbf_ctx* c : b->ctx) {
    const size_t bytes[3] = {10 * c->NS * c->tsz, 3 * c->NS * c->tsz, c->p.twocomponent ? c->N * c->asz : 0};
    void* dst[3] = {c->h[c->cur], c->U[c->ucur], c->p.twocomponent ? c->phi[c->cur] : nullptr};
    for (int i = 0; i < 3; ++i, ++k)
      if (bytes[i]) BHIPCHK(b, hipMemcpyAsync(dst[i], snap[k], bytes[i], hipMemcpyDeviceToDevice, b->stream));
  }
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  release();
  int rc = batch_tables(b);
  if (rc) return rc;
  if (best_ty) *best_ty = b->ty;
  if (best_tz) *best_tz = btz;
  b->timing_valid = false;
  return LBF_OK;
}

int lbf_batch_set_order(lbf_batch* b, const int32_t* order) {
  if (!b || !order) return LBF_EINVAL;
  const int nd = (int)b->ctx.size();
  std::vector<int> o(order, order + nd), seen(nd, 0);
  for (int v : o) {
    if (v < 0 || v >= nd || seen[v]) return bfail(b, LBF_EINVAL, "order must be a permutation of 0..%d", nd - 1);
    seen[v] = 1;
  }
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  {
    int rc = batch_refresh(b);
    if (rc) return rc;
  }
  batch_drop_graph(b);
  b->order = o;
  return batch_tables(b);
}

int lbf_batch_tiles(const lbf_batch* b, int32_t* tile_y, int32_t* tile_z, int32_t* tile_z_prep, int32_t* order) {
  if (!b) return LBF_EINVAL;
  const int nd = (int)b->ctx.size();
  if (tile_y) *tile_y = b->ty;
  for (int d = 0; d < nd; ++d) {
    if (tile_z) tile_z[d] = b->tz[d];
    if (tile_z_prep) tile_z_prep[d] = b->tz_prep[d];
    if (order) order[d] = b->order[d];
  }
  return LBF_OK;
}

}  // extern "C"
