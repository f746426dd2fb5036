/*
 * TEST INFRASTRUCTURE ONLY — the oracle.  Nothing in the product links, loads
 * or calls this code; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it, and only as the checker.
 *
 * CPU restatement of the reference DPM kernels, shaders/Cell2D_kernel.cl and
 * shaders/Cell3D_Kernel.cl, driven as Tissue2D.cpp:116-241 and
 * Tissue3D.cpp:118-522 drive them.  Two instantiations of one body
 * (dpm_oracle_tmpl.h): *_f32 keeps the reference's float arithmetic with its
 * double-literal promotions, *_f64 evaluates the same formulas in double.
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off, no -ffast-math).
 * An OpenMP build of the same source (-fopenmp) only parallelises the outer
 * per-cell loops of ora3d_forces for the CPU baseline timing; results are
 * identical because every output element is written by one iteration.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#if defined(ORA_OMP)
#define ORA_PARALLEL_FOR _Pragma("omp parallel for schedule(dynamic)")
#else
#define ORA_PARALLEL_FOR
#endif

/* ---- float instantiation ("ref32") ---- */
#define R float
#define FN(name) name##_f32
#define SQRT sqrtf
#define FABS fabsf
#define ROUND roundf
#define FLOOR floorf
#define ATAN2 atan2f
#define FMA fmaf
#include "dpm_oracle_tmpl.h"
#undef R
#undef FN
#undef SQRT
#undef FABS
#undef ROUND
#undef FLOOR
#undef ATAN2
#undef FMA

/* ---- double instantiation ("ref64") ---- */
#define R double
#define FN(name) name##_f64
#define SQRT sqrt
#define FABS fabs
#define ROUND round
#define FLOOR floor
#define ATAN2 atan2
#define FMA fma
#include "dpm_oracle_tmpl.h"
#undef R
#undef FN
#undef SQRT
#undef FABS
#undef ROUND
#undef FLOOR
#undef ATAN2
#undef FMA

int ora_abi_version(void) { return 1; }
