/*
 * lbfast_hip.h -- C-ABI of the MI355X (gfx950) moment-represented lattice
 * Boltzmann time step.  Plain pointers and sizes only: a Fortran host binds
 * it with iso_c_binding, Python with ctypes.
 *
 * Every entry point returns 0 on success and a non-zero LBF_E* code on
 * failure; lbf_last_error() then holds a message.  A context owns all of
 * its device memory and one HIP stream; host buffers are borrowed for the
 * duration of a call.  Calls on one context must come from one host thread.
 *
 * Field arrays crossing the boundary are interior-only, Fortran order
 * (i fastest, then j, then k), nx*ny*nz values -- the order of the
 * reference's RAW files (print_module.f90:826-848).  The 1-node halo of the
 * reference (allocate_module.f90:19-55) does not exist here: every kernel
 * wraps indices periodically (mpi_module.f90:2820-2918 semantics).
 *
 * Reference interface replaced (lb_cuda_driver_module.f90, called from
 * integrator_module.f90:266-827 and boundary_cds_module.f90:49-56,1345-1351):
 *   setup_cuda                      (driver:44-168)    -> lbf_create
 *   compute_norm_interface_cuda     (driver:545-819)   -> lbf_stage(LBF_STAGE_NORM)
 *   compute_div_theta_n             (driver:821-848)   -> lbf_stage(LBF_STAGE_DIV_THETA_N)
 *   moments_LB_cuda                 (driver:912-1430)  -> lbf_stage(LBF_STAGE_MOMENTS)
 *   fused_LB_cuda                   (driver:1432-1518) -> lbf_stage(LBF_STAGE_FUSED)
 *   update_phifields                (driver:1520-1899) -> lbf_stage(LBF_STAGE_UPDATE_PHI)
 *   LB_int_boundary_cuda            (driver:1901-1963) -> lbf_stage(LBF_STAGE_LB_INT_BOUNDARY)
 *   PHI_int_boundary_cuda           (driver:1965-2009) -> lbf_stage(LBF_STAGE_PHI_INT_BOUNDARY)
 *   thinfilm_scan_mark_cuda         (driver:850-880)   -> lbf_stage(LBF_STAGE_THINFILM_SCAN)
 *   repulsive_flux_normal_cuda      (driver:882-910)   -> lbf_stage(LBF_STAGE_REPULSIVE_FLUX)
 *   phi_sum_count_cuda + apply_lagrangian_phi_cuda (driver:2011-2077, BCPHIFLUX)
 *                                                      -> lbf_stage(LBF_STAGE_PHI_LAGRANGE)
 *   open-boundary loops of bcs_mesoscopic_hfields / _phifields
 *     (boundary_cds_module.f90:57-1311, 1353-1664)     -> lbf_stage(LBF_STAGE_OPEN_BC_H / _PHI)
 *   fused_LB_cuda under BOUNCE_BACK (driver:1464-1465, fused_LB_kernel_bb)
 *                                                      -> lbf_stage(LBF_STAGE_FUSED), bounce_back = 1
 *   exchange_*_intpbc (mpi_module.f90:2203-3524)       -> in-kernel periodic wrap (no call)
 *   integrator time loop (integrator_module.f90:266-827) -> lbf_step (fused fast path)
 *   initial O:=S copy (init_conditions_module.f90:584-587) + bcs_phi
 *                     (integrator_module.f90:179-182)  -> lbf_init_state
 *   isfluid classification (LBFAST.f90:258-328)         -> lbf_classify_isfluid
 *   copy_print observables (print_module.f90:638-776)   -> lbf_download_field(LBF_FIELD_RHOPHI ...)
 */
#ifndef LBFAST_HIP_H
#define LBFAST_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBF_ABI_VERSION 2

/* ---- status codes ------------------------------------------------------ */
#define LBF_OK          0
#define LBF_EINVAL      1   /* bad argument / unsupported combination        */
#define LBF_ESIZE       2   /* domain size rejected (setup_cuda checks)        */
#define LBF_EHIP        3   /* HIP runtime error (launch, alloc, copy)         */
#define LBF_ENODEV      4   /* no gfx950 device visible                        */
#define LBF_ESTATE      5   /* call out of order (e.g. step before init)       */

/* ---- precision modes (defines.h PRC / STRPRC / MIXEDPRC) --------------- */
#define LBF_PREC_FP64     0  /* PRC 8, STRPRC 8                              */
#define LBF_PREC_FP32     1  /* PRC 4, STRPRC 4                              */
#define LBF_PREC_MIXED    2  /* PRC 8 compute, STRPRC 4 storage ("sd")       */
#define LBF_PREC_FP16S    3  /* FP32 compute, moments and u_old stored as
                                scaled FP16 deviations (P, u, Pi - u u), the
                                time step rounding stochastically; phi and the
                                phase-field intermediates FP32 */
#define LBF_PREC_BF16S    4  /* as FP16S with bfloat16 storage               */

/* ---- field ids ----------------------------------------------------------
 * 1..10 follow the reference hfields order (vars_module.f90:731):
 * P (density deviation), u, v, w, Pxx, Pyy, Pzz, Pxy, Pxz, Pyz.          */
#define LBF_FIELD_P        1
#define LBF_FIELD_U        2
#define LBF_FIELD_V        3
#define LBF_FIELD_W        4
#define LBF_FIELD_PXX      5
#define LBF_FIELD_PYY      6
#define LBF_FIELD_PZZ      7
#define LBF_FIELD_PXY      8
#define LBF_FIELD_PXZ      9
#define LBF_FIELD_PYZ     10
#define LBF_FIELD_PHI     11  /* phase field of the current (S) buffer        */
#define LBF_FIELD_UOLD_X  12  /* previous step's corrected velocity (the      */
#define LBF_FIELD_UOLD_Y  13  /*  reference reads it from the other hfields   */
#define LBF_FIELD_UOLD_Z  14  /*  buffer, lb_cuda_moments_module.f90:192-194) */
#define LBF_FIELD_FX      15  /* forces (stage path only)                     */
#define LBF_FIELD_FY      16
#define LBF_FIELD_FZ      17
#define LBF_FIELD_AUX1    18  /* aux 1..7: nx ny nz |grad phi| theta_x..z      */
#define LBF_FIELD_AUX7    24  /*  (lb_cuda_auxfields_module.f90:104-118)       */
#define LBF_FIELD_LAPPHI  25  /* locaux 1: laplacian of phi                    */
#define LBF_FIELD_DIVTHN  26  /* locaux 2: div(phi(1-phi) n)                   */
#define LBF_FIELD_P_O     27  /* the other (O) moments buffer, ids 27..36       */
#define LBF_FIELD_PHI_O   37  /* the other (O) phase-field buffer              */
#define LBF_FIELD_RHOPHI  40  /* observable rho_r phi + (1-phi) rho_b (download only;
                                 print_module.f90:734-760)                     */
#define LBF_FIELD_LOCAUX3 41  /* locaux 3..5: facing-partner indices (1-based,  */
#define LBF_FIELD_LOCAUX5 43  /*  unwrapped; lb_cuda_repulsive_module.f90:144-146) */
#define LBF_FIELD_LOCAUX6 44  /* locaux 6..8: repulsive flux (repulsive:285-287) */
#define LBF_FIELD_LOCAUX8 46
#define LBF_FIELD_REP_MASK 47 /* rep_mask (0/1, download only)                 */
#define LBF_NFIELD_IDS    48

/* ---- stage ids (lbf_stage) --------------------------------------------- */
#define LBF_STAGE_NORM              1  /* compute_norm_interface_kernel    */
#define LBF_STAGE_PHI_INT_BOUNDARY  2  /* PHI_int_boundary_kernel          */
#define LBF_STAGE_DIV_THETA_N       3  /* compute_div_theta_n_kernel       */
#define LBF_STAGE_MOMENTS           4  /* moments_LB_kernel                */
#define LBF_STAGE_FUSED             5  /* fused_LB_kernel                  */
#define LBF_STAGE_UPDATE_PHI        6  /* update_phifields_kernel          */
#define LBF_STAGE_LB_INT_BOUNDARY   7  /* LB_int_boundary_kernel           */
#define LBF_STAGE_THINFILM_SCAN     8  /* thinfilm_scan_mark_kernel (REPULSIVE_FLUX) */
#define LBF_STAGE_REPULSIVE_FLUX    9  /* repulsive_flux_normal_kernel     */
#define LBF_STAGE_PHI_LAGRANGE     10  /* phi_sum_count_kernel + apply_lagrangian_phi_kernel
                                          (BCPHIFLUX), after PHI_INT_BOUNDARY (and OPEN_BC_PHI);
                                          with openbc the target adds the phase flux
                                          through the open faces' first fluid layers */
#define LBF_STAGE_OPEN_BC_PHI      11  /* open faces, phase field: the wall layer takes
                                          phi of the node inside it
                                          (boundary_cds_module.f90:1353-1664) */
#define LBF_STAGE_OPEN_BC_H        12  /* open faces, moments: after LB_INT_BOUNDARY, the
                                          populations a first-layer node receives from
                                          solid nodes are replaced by the regularised
                                          populations of the face's pressure / velocity
                                          (boundary_cds_module.f90:57-1311) */

/* Parameters: the &simulation / &fluid namelists (LBFAST.f90:32-60) that
 * the time step reads, the defines.h switches, and the derived quantities
 * of LBFAST.f90:212-219 (filled by lbf_params_derive). */
typedef struct lbf_params {
  int32_t nx, ny, nz;            /* lx, ly, lz (one rank)                   */
  int32_t lattice;               /* LATTICE: 19 or 27                       */
  int32_t highorder;             /* HIGHORDER (LATTICE 27 only)             */
  int32_t twocomponent;          /* TWOCOMPONENT                            */
  int32_t densratio;             /* DENSRATIO (requires TWOCOMPONENT)       */
  int32_t interface_incomp;      /* INTERFACE_INCOMP (defines.h:72: implied by DENSRATIO) */
  int32_t internal_obstacles;    /* INTERNAL_OBSTACLES                      */
  int32_t buoyancy_forcing;      /* BUOYANCY_FORCING                        */
  int32_t pbc[3];                /* pbc_x pbc_y pbc_z (1 periodic, 0 wall)  */
  int32_t precision;             /* LBF_PREC_*                              */
  int32_t bounce_back;           /* BOUNCE_BACK (with internal_obstacles)    */
  int32_t smagorinsky;           /* SMAGORINSKI                             */
  int32_t repulsive_flux;        /* REPULSIVE_FLUX (two-component)          */
  int32_t bcphiflux;             /* BCPHIFLUX: global phase-mass correction
                                    (two-component with internal obstacles) */
  double fx, fy, fz;             /* body force                              */
  double visc1, visc2;           /* kinematic viscosities                   */
  double rho_r, rho_b;           /* phase densities (DENSRATIO)             */
  double tau_diff, sigma, width; /* phase-field mobility, tension, width    */
  double wettab_r;               /* contact angle in degrees (90 default)   */
  /* derived by lbf_params_derive (LBFAST.f90:212-219) */
  double tau1, omega, tau2, beta, kapp, sharp_c;
  /* FP16S/BF16S: stored value = half(scale * deviation); 0 selects default */
  double store_scale[10];
  /* REPULSIVE_FLUX namelist values (LBFAST.f90:57); the scan window is 3 (LBFAST.f90:221) */
  double a_rep, q_th, cos_opp_t;
  double reserved_d;
  /* open boundaries (&simulation, LBFAST.f90:33-37; internal obstacles only):
   * openbc = 1 applies them on every non-periodic face.  Faces in the order
   * x(1), x(2), y(1), y(2), z(1), z(2) (low, high).  openbc_type: 2 the
   * node's own pressure and velocity, 1 the face's openbc_press with the
   * node's velocity, 0 the node's pressure with the face's velocity
   * (openbc_u, openbc_v, openbc_w). */
  int32_t openbc;
  int32_t openbc_type[6];
  int32_t reserved_i2;
  double openbc_press[6], openbc_u[6], openbc_v[6], openbc_w[6];
} lbf_params;

typedef struct lbf_ctx lbf_ctx;

/* Fill tau1, omega, tau2, beta, kapp, sharp_c from the namelist values,
 * exactly as LBFAST.f90:212-219. Pure host function. */
int lbf_params_derive(lbf_params* p);

/* Set defaults: periodic, D3Q27 HIGHORDER, FP32, 1c, all physics zero. */
void lbf_params_default(lbf_params* p);

/* Classify nodes as LBFAST.f90:258-328 does: start from `in` (1 fluid, 0 solid,
 * e.g. an isfluid.raw file), put walls at non-periodic global faces when
 * internal_obstacles is set, then mark fluid nodes with a non-fluid
 * neighbour on the compiled lattice as -1.  Pure host function.
 * in/out: nx*ny*nz int8, Fortran order (in may equal out). */
int lbf_classify_isfluid(const lbf_params* p, const int8_t* in, int8_t* out);

/* Create a context on the current HIP device (setup_cuda + allocate_struct).
 * Rejects sizes the kernels cannot tile and unsupported switch sets. */
int lbf_create(const lbf_params* p, lbf_ctx** out);
void lbf_destroy(lbf_ctx* ctx);
const char* lbf_last_error(const lbf_ctx* ctx);   /* ctx may be NULL */
const char* lbf_global_last_error(void);          /* errors of lbf_create */

/* dtype of host arrays: 0 = float64, 1 = float32 */
int lbf_upload_field(lbf_ctx* ctx, int field_id, const void* host, int dtype);
int lbf_download_field(lbf_ctx* ctx, int field_id, void* host, int dtype);
/* Classified isfluid (1, 0, -1), nx*ny*nz int8 Fortran order. */
int lbf_upload_isfluid(lbf_ctx* ctx, const int8_t* host);

/* After uploading P..Pyz (and phi): O := S, phi_O := phi_S, u_old := u,
 * aux := 0 (init_conditions_module.f90:177-220,584-587), then the initial
 * bcs_phi ghost update (integrator_module.f90:179-182). */
int lbf_init_state(lbf_ctx* ctx);

/* Advance nsteps full time steps with the fused fast path (stream-ordered;
 * returns after enqueueing).  Exactly nsteps: a host reproducing the
 * reference loop's pairing (integrator_module.f90:266) passes 2*ceil(n/2).
 * With open faces each step adds the open-face phase copy before and the
 * open-face correction after the fused launch; REPULSIVE_FLUX runs the
 * stage sequence. */
int lbf_step(lbf_ctx* ctx, int32_t nsteps);

/* Stage path: the reference's per-kernel entry points, in the reference's
 * buffer semantics (moments in place on S, u_old from the previous step).
 * lbf_begin_step advances the step counter; lbf_end_step swaps S and O. */
int lbf_begin_step(lbf_ctx* ctx);
int lbf_stage(lbf_ctx* ctx, int stage_id);
int lbf_end_step(lbf_ctx* ctx);

/* FP64 reductions over interior nodes with |isfluid| == 1 of the current
 * state: out[0] = sum P, out[1..3] = sum rhophi*u (momentum, rhophi = 1 in
 * 1c), out[4] = sum phi, out[5] = sum rhophi (mass), out[6] = sum u.u,
 * out[7] = fluid node count.  Synchronises. */
int lbf_reduce_fp64(lbf_ctx* ctx, double out[8]);

/* lbf_reduce_fp64 without allocation.  No counterpart in the reference: it
 * is for callers that cannot call hipMalloc (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation).  The partial sums
 * go to a device array of the library (1024 records of 8 doubles plus the 8
 * sums, 64 KiB + 64 B, held by each GPU the library runs on until the
 * process ends) instead of the context's buffers; the grid-stride reduction
 * covers every domain size lbf_create accepts within those 1024 records, so
 * no size is refused.  The results equal lbf_reduce_fp64's bit for bit.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (from several host threads or contexts) race.  Synchronises. */
int lbf_reduce_fp64_noalloc(lbf_ctx* ctx, double out[8]);

int lbf_synchronize(lbf_ctx* ctx);
int32_t lbf_get_step(const lbf_ctx* ctx);

/* Device pointer of a state field (for zero-copy interop); NULL if absent.
 * Storage type follows the precision mode.  Single fields (phi, lap phi,
 * div theta n) are x-fastest nx*ny*nz arrays.  A field of a field set (the
 * ten moments, u_old, forces, aux, repulsive locaux) is row-interleaved with
 * the other F fields of its set: the pointer is its first row, element
 * (i, j, k) (0-based) is at index (k*ny + j)*F*sx + i, with sx = nx rounded
 * up to even and F = 10, 3, 3, 7, 6 respectively. */
void* lbf_field_device_ptr(lbf_ctx* ctx, int field_id);

/* Timing of the last lbf_step call, measured with HIP events on the context
 * stream: total ms of the call, and the summed ms of its fused-kernel
 * launches when per-launch timing is on (else fused_ms = -1).  No
 * counterpart in the reference beyond its host timers (profiling_m.f90). */
int lbf_last_step_timing(lbf_ctx* ctx, double* total_ms, double* fused_ms);
/* Per-launch timing (off by default): when on, every fused launch of
 * lbf_step / lbf_stage is bracketed by its own HIP event pair, which costs
 * two event records per step and one event pair per step of the longest
 * call.  Profilers and benchmarks turn it on; long production runs leave it off. */
int lbf_set_timing(lbf_ctx* ctx, int32_t on);

/* Fused-kernel tile shape override (0 = default/autotuned). */
int lbf_set_tile(lbf_ctx* ctx, int32_t tile_y, int32_t tile_z);

int32_t lbf_abi_version(void);

/* ---- batches of independent domains ----------------------------------------
 * No counterpart in the reference, which runs one domain per process.  Many
 * independent domains (initialised contexts sharing lattice, HIGHORDER,
 * precision and component count; any sizes and physical parameters; no
 * obstacles or REPULSIVE_FLUX) advance together: each step is one phase-field
 * launch and one fused launch whose grids cover every domain's tiles, so that
 * small domains fill the GPU.  The batch borrows the contexts (they must
 * outlive it) and keeps their step counters and buffers current, so uploads,
 * downloads and single-domain calls stay valid between batch calls; work on
 * a context's stream is ordered before and after each batch step.  Two-
 * component batches need float phase-field storage (not FP64). */
typedef struct lbf_batch lbf_batch;
int lbf_batch_create(lbf_ctx* const* ctxs, int32_t ndomains, lbf_batch** out);
void lbf_batch_destroy(lbf_batch* b);
const char* lbf_batch_last_error(const lbf_batch* b);
/* Advance every domain nsteps (stream-ordered; replays a captured graph for
 * whole multiples of its length, launches the rest directly). */
int lbf_batch_step(lbf_batch* b, int32_t nsteps);
/* Capture nsteps (even, >= 2) steps into a hipGraph that later
 * lbf_batch_step calls replay; set_tile and autotune discard it. */
int lbf_batch_capture(lbf_batch* b, int32_t nsteps);
/* Row tile (8 or 13) of the fused kernel and z chunks of the fused and
 * phase-field kernels, for every domain (chunks clipped to each nz). */
int lbf_batch_set_tile(lbf_batch* b, int32_t tile_y, int32_t tile_z, int32_t tile_z_prep);
/* The table order: order[i] is the domain whose tiles come i-th in the
 * grid of every launch (each XCD takes a contiguous eighth of it); a
 * permutation of 0..ndomains-1.  Results do not depend on it. */
int lbf_batch_set_order(lbf_batch* b, const int32_t* order);
/* Time trial_steps (even) steps of each candidate on the batch itself, keep
 * the fastest and restore every domain's state.  Searched in turn: the row
 * tile with one z chunk for all domains (reported in tile_y, tile_z); the z
 * chunk of each size class (domains of one nz); the table order (as given,
 * largest first, largest and smallest alternating). */
int lbf_batch_autotune(lbf_batch* b, int32_t trial_steps, int32_t* tile_y, int32_t* tile_z);
/* The batch's current choices: row tile, and per domain (arrays of
 * ndomains) the fused and phase-field z chunks and the table order. */
int lbf_batch_tiles(const lbf_batch* b, int32_t* tile_y, int32_t* tile_z, int32_t* tile_z_prep, int32_t* order);
int lbf_batch_synchronize(lbf_batch* b);
/* Elapsed milliseconds of the last lbf_batch_step call (HIP events). */
int lbf_batch_timing(lbf_batch* b, double* ms);

#ifdef __cplusplus
}
#endif
#endif /* LBFAST_HIP_H */
