/* lbfast_defines.h -- compile-time switches of the Fortran host, with the
 * implications among them that the reference's defines.h (lines 38-74)
 * states.  Select with -D on the amdflang line, e.g.
 *   -DLAPLACE  -DLATTICE=27 -DHIGHORDER  -DPRC=4
 * LATTICE (19|27), HIGHORDER, TWOCOMPONENT, DENSRATIO, INTERNAL_OBSTACLES,
 * BUOYANCY_FORCING, BOUNCE_BACK, SMAGORINSKI, REPULSIVE_FLUX, BCPHIFLUX (the
 * global phase-mass correction; two-component runs with obstacles); precision:
 * PRC (8|4), MIXEDPRC (PRC 8 arithmetic, 4-byte storage), STORE_FP16 /
 * STORE_BF16 (16-bit scaled-deviation storage, FP32 arithmetic); test case:
 * TAYLORGREEN, LAPLACE, LAMBTEST, POISEUILLE, TWOPOISEUILLE (default: the
 * droplet of the generic initial condition); output: WRITEPRESS, PRINTPHI,
 * DOXDMF (an XDMF description beside every RAW frame).
 * DOBENCHMARK (on, as in the reference; -DNO_DOBENCHMARK turns it off) sets
 * the benchmark defaults that an input deck then overrides: a MYSIDE^3 box
 * (MYSIDE 512), 100 steps, RAW output every 100 steps. */
#ifndef NO_DOBENCHMARK
#define DOBENCHMARK
#endif
#ifndef MYSIDE
#define MYSIDE 512
#endif
#ifndef LATTICE
#define LATTICE 27
#endif
#ifndef PRC
#define PRC 8
#endif
#if defined(LAPLACE) || defined(LAMBTEST)
#define TWOCOMPONENT
#define DENSRATIO
#define PRINTPHI
#define WRITEPRESS
#endif
#ifdef TWOPOISEUILLE
#define INTERNAL_OBSTACLES
#define TWOCOMPONENT
#define DENSRATIO
#define PRINTPHI
#endif
#ifdef POISEUILLE
#define INTERNAL_OBSTACLES
#undef TWOCOMPONENT
#undef DENSRATIO
#endif
#ifdef TAYLORGREEN
#undef TWOCOMPONENT
#undef DENSRATIO
#endif
#ifdef DENSRATIO
#define INTERFACE_INCOMP
#endif
