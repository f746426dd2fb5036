/*
 * dpm.h — C ABI of the MI355X (gfx950) Deformable Particle Model timestep.
 *
 * The reference (sudo-shaka/OpenCL_DPM) runs its timestep inside the C++
 * methods Tissue2D::CLEulerUpdate (src/Tissue2D.cpp:116-241) and
 * Tissue3D::CLEulerUpdate (src/Tissue3D.cpp:118-522), which pack the cells,
 * JIT-build shaders/Cell2D_kernel.cl / Cell3D_Kernel.cl, upload, launch a
 * chain of kernels per step and read back.  This library replaces that
 * body: a tissue is created once as an opaque handle with device-resident
 * state, positions are set and read through plain host arrays, and
 * dpm_euler() advances it.  No HIP, OpenCL or torch type crosses the ABI.
 *
 * Conventions: every function returns DPM_OK (0) or a DPM_E* status; the
 * message of the last failure on the calling thread is dpm_last_error().
 * DPM_EINVAL corresponds to the reference's std::invalid_argument
 * (Tissue3D.cpp:123-135, Python ValueError); DPM_ENONFINITE, DPM_EDEVICE
 * and DPM_ERANGE to its std::runtime_error (Python RuntimeError).
 * A handle is not thread-safe; distinct handles are independent.
 */
#ifndef DPM_H_
#define DPM_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DPM_OK 0
#define DPM_EINVAL 1      /* bad argument (ValueError) */
#define DPM_ENONFINITE 2  /* non-finite position (RuntimeError) */
#define DPM_EDEVICE 3     /* HIP runtime failure or no device (RuntimeError) */
#define DPM_ENOMEM 4      /* device allocation failed (RuntimeError) */
#define DPM_ERANGE 5      /* size beyond a documented limit (RuntimeError) */

typedef struct dpm_tissue dpm_tissue;

/* Force-term masks.  The 2D bits follow the reference kernel order
 * (Tissue2D.cpp:216-222); the 3D bits the launch order of
 * Tissue3D.cpp:374-414 plus the two terms the reference does not launch. */
#define DPM2D_AREA 1u
#define DPM2D_PERIMETER 2u
#define DPM2D_BENDING 4u
#define DPM2D_ATTRACTION 8u
#define DPM2D_REPULSION 16u
#define DPM2D_ALL 31u       /* exactly the terms Tissue2D::CLEulerUpdate runs */
/* Two 2D terms with no reference counterpart that make the force law
 * conservative (the reference's area force has a non-gradient y sign and
 * its crossing repulsion is not a gradient), for FIRE and jamming:
 * AREA_GRAD = -grad of Ka sqrt(a0)/2 (|A|/a0 - 1)^2 (equals the reference's
 * x component on counter-clockwise cells); VREPULSION = soft disks on the
 * vertices of different cells, U = Kre/2 (1 - d/sigma)^2 for
 * d < sigma = (l0_i + l0_j)/2, minimum image. */
#define DPM2D_AREA_GRAD 32u
#define DPM2D_VREPULSION 64u
#define DPM2D_CONSERVATIVE 102u /* AREA_GRAD | PERIMETER | BENDING | VREPULSION */
#define DPM2D_ALL_TERMS 127u

#define DPM3D_VOLUME 1u     /* VolumeForceUpdate, Cell3D_Kernel.cl:66-112 */
#define DPM3D_SURFACE 2u    /* SurfaceAreaForceUpdate, :114-177 */
#define DPM3D_SUBSTRATE 4u  /* StickToSurface, :180-247 */
#define DPM3D_CONTACT 8u    /* RepellingForces, :251-310 */
#define DPM3D_ADHESION 16u  /* AllVertAttraction, :313-364 (never launched by
                               the reference host; opt-in here) */
#define DPM3D_BENDING 32u   /* graph-Laplacian bending, no reference kernel */
#define DPM3D_REFERENCE 15u /* exactly the terms Tissue3D::CLEulerUpdate runs */
#define DPM3D_ALL 63u

const char* dpm_last_error(void);
int dpm_version(void);
/* Number of visible HIP devices (0 on a machine without one). */
int dpm_device_count(int* n);

/* ------------------------------------------------------------------ 3D ---
 * A batch of `ntissues` independent tissues (1 for the reference's single
 * Tissue3D) that share one closed triangle mesh topology: every cell has nv
 * vertices and nf faces (nv = 10*4^s + 2, nf = 20*4^s for an icosphere of
 * subdivision s; the reference fixes s = 2, cell.hpp:30-31).  Cells of all
 * tissues are numbered consecutively; tissue t owns cells
 * [cell_begin[t], cell_begin[t+1]).
 * Replaces the packing + buffer set-up of Tissue3D::CLEulerUpdate,
 * Tissue3D.cpp:139-281; per-cell l0 = sqrt(4 a0)/sqrt(3) as Tissue3D.cpp:177.
 */
typedef struct {
  int ntissues;
  const int* cell_begin;  /* [ntissues+1] */
  int nv, nf;
  const int* faces;       /* [nf*3] vertex indices, outward orientation */
  /* per cell [ncells], ncells = cell_begin[ntissues] */
  const float* Kv;
  const float* Ka;
  const float* Ks;
  const float* Kb;        /* may be NULL (all 0) */
  const float* v0;
  const float* a0;
  /* per tissue [ntissues] */
  const float* L;
  const int* pbc;
  const float* Kre;       /* contact stiffness ("Kc" of RepellingForces) */
  const float* Kat;       /* vertex-vertex adhesion stiffness */
  const float* ks_grad;   /* substrate stiffness Ks*(1 + ks_grad*x/L); NULL = 0 */
  unsigned mask;          /* DPM3D_* terms */
} dpm3d_desc;

/* Creates the handle and allocates device state on the current device.
 * Validates as Tissue3D.cpp:145-170 does (face indices, Kv, Ka, Ks > 0,
 * v0, a0 > 0). */
int dpm3d_create(const dpm3d_desc* desc, dpm_tissue** out);
/* Positions in the reference's per-cell AoS order, xyz[(c*nv + v)*3 + k];
 * n = ncells*nv.  set_positions rejects non-finite input
 * (Tissue3D.cpp:184-190). */
int dpm3d_set_positions(dpm_tissue* h, const float* xyz, size_t n);
int dpm3d_get_positions(dpm_tissue* h, float* xyz, size_t n);
/* Forces of the last step taken (Tissue3D.cpp:425-434, 498-499). */
int dpm3d_get_forces(dpm_tissue* h, float* xyz, size_t n);
/* Per-cell and per-tissue parameters after creation (any pointer may be
 * NULL = unchanged; the mask is always replaced).  The reference re-reads
 * them from the cells on every call. */
int dpm3d_set_cell_params(dpm_tissue* h, const float* Kv, const float* Ka,
                          const float* Ks, const float* Kb, const float* v0,
                          const float* a0);
int dpm3d_set_tissue_params(dpm_tissue* h, const float* L, const int* pbc,
                            const float* Kre, const float* Kat,
                            const float* ks_grad, unsigned mask);
/* Per-cell volume of the current positions (double) and the centre of mass
 * (float xyz) — the host reductions GetVolume/GetCOM, cell.cpp:196-211,
 * 276-288, computed on the device.  Either pointer may be NULL. */
int dpm3d_cell_stats(dpm_tissue* h, double* volume, float* com);
/* Diagnostic: evaluates the forces once at the current positions without
 * moving them, writes them to `forces` (AoS) and records the pair sets the
 * reference predicates select: contact[(gv)*ncmax + (cj - first cell of the
 * tissue)] = 1 where vertex gv and cell cj are in contact (|w| >= 1e-6,
 * Cell3D_Kernel.cl:303), ncmax = largest tissue's cell count; winding (may
 * be NULL) the winding number of every evaluated pair (0 where culled);
 * down[c*nf + f] = 1 where face f of cell c faces the substrate
 * (Cell3D_Kernel.cl:214); adhesion[gv] = number of adhesion partners of
 * vertex gv on its own side (Cell3D_Kernel.cl:353).  Any output may be NULL. */
int dpm3d_probe(dpm_tissue* h, float* forces, unsigned char* contact,
                float* winding, unsigned char* down, int* adhesion);

/* Diagnostic: the adhesion pair sets of one force evaluation at the current
 * positions (as dpm3d_probe, which gives their sizes).  A partner vertex is
 * identified by the bit patterns of its float coordinates,
 * key = f(x ^ f(y ^ f(z))) with f the 32-bit finaliser of MurmurHash3;
 * per vertex gv, own_keys[gv] is the wrapping 32-bit sum of the keys of the
 * partners its own loop accepts (Cell3D_Kernel.cl:341-353, |d| < 2 l0 of
 * gv's cell) and rev_keys[gv] that of the vertices whose own loop accepts gv
 * (the pairs whose reaction lands on gv, Cell3D_Kernel.cl:359).  Either may
 * be NULL. */
int dpm3d_probe_adhesion(dpm_tissue* h, unsigned* own_keys, unsigned* rev_keys);
/* Diagnostic: the contact pair set of one force evaluation in compact form,
 * for tissues too large for dpm3d_probe's vertex x cell table: per vertex
 * the number of cells it is in contact with (|w| >= 1e-6,
 * Cell3D_Kernel.cl:303) and the wrapping 32-bit sum of those cells' keys
 * f(c + 0x9E3779B9), c the cell's index within its tissue and f the
 * finaliser of dpm3d_probe_adhesion.  Either may be NULL. */
int dpm3d_probe_contact(dpm_tissue* h, unsigned* count, unsigned* keys);

/* dpm3d_probe without device allocation (no reference counterpart; for
 * callers that cannot allocate device memory: tight serving loops, harnesses
 * that forbid hipMalloc).  Same outputs as dpm3d_probe, bit for bit; the
 * device tables live in a 64 MiB array of the library that every device
 * holds from the library's load until the process ends.  Fails with
 * DPM_ERANGE, before any device work, when the requested tables (contact and
 * winding nverts * ncmax bytes and floats, down ncells * nf bytes, adhesion
 * nverts ints, forces 3 * nverts floats, each rounded up to 256 bytes)
 * exceed 64 MiB.  Every call on a device shares that device's array: calls
 * that overlap on one device (other threads, other handles) race. */
int dpm3d_probe_noalloc(dpm_tissue* h, float* forces, unsigned char* contact,
                        float* winding, unsigned char* down, int* adhesion);

/* The reference's host reductions Cell3D::GetVolume, GetSurfaceArea and
 * GetCOM (cell.cpp:196-230, 276-288) of every cell at the current
 * positions, evaluated on the device in the host functions' float operation
 * order (bit-identical to them; Tissue3D::CLEulerUpdate recomputes volume and
 * surface area on the host after every call, Tissue3D.cpp:503-504).  The
 * surface area is the host function's: the sum of |cross| over faces, twice
 * the triangulated area.  volume[ncells], surface_area[ncells],
 * com[ncells*3]; any may be NULL. */
int dpm3d_cell_measures(dpm_tissue* h, float* volume, float* surface_area, float* com);

/* Full double precision.  dpm3d_create_f64 makes a handle whose positions,
 * forces and per-cell statistics are double on the device and whose step
 * evaluates every formula in double (the float thresholds of the reference
 * kernels promoted); the reference's kernels are float only
 * (Cell3D_Kernel.cl), so this is the FP64 comparison run of a precision
 * study, not a reference entry point.  All dpm3d_* and dpm_* calls accept
 * either kind of handle; the _f64 accessors move doubles without rounding
 * (on a float handle they convert exactly from float). */
int dpm3d_create_f64(const dpm3d_desc* desc, dpm_tissue** out);
int dpm3d_set_positions_f64(dpm_tissue* h, const double* xyz, size_t n);
int dpm3d_get_positions_f64(dpm_tissue* h, double* xyz, size_t n);
int dpm3d_get_forces_f64(dpm_tissue* h, double* xyz, size_t n);
/* Elastic energy of each cell at the current positions, in double:
 *   Kv v0/2 (V/v0 - 1)^2 + sum over face edges sc l0/2 (l/l0 - 1)^2
 *   + Kb/2 sum_v |sum_{u~v} (x_u - x_v)|^2,   sc = 0.3 Ka sqrt(a0)/l0,
 * restricted to the terms the mask enables; minus its gradient is the
 * volume + surface + bending force of the step.  The reference has no
 * energy (the substrate, contact and adhesion forces are not conservative);
 * it is the drift diagnostic of the precision study.  energy[ncells]. */
int dpm3d_energy(dpm_tissue* h, double* energy);

/* Candidate cells of the contact/adhesion search: 0 scans every cell of the
 * tissue (the reference's all-pairs loops, Cell3D_Kernel.cl:273-310,
 * 325-364, with exact culling); 1 bins the centres of mass each step
 * (counting sort into hashed periodic bins no narrower than the step's
 * largest candidate reach) and visits the 27 surrounding bins; -1 (default)
 * chooses 1 for tissues of more than 8 x threads-per-cell cells.  Both give
 * bit-identical results. */
int dpm3d_set_cell_list(dpm_tissue* h, int mode);

/* Bin width of the cell list over the step's largest candidate reach, a
 * factor in [1, 4] (default 1, the narrowest bins the 27-bucket gather
 * allows); wider bins mean fewer, fuller buckets.  Every factor gives
 * bit-identical results.  No reference counterpart (the reference scans
 * every cell pair, Cell3D_Kernel.cl:273-310, 325-364); dpm3d_autotune tunes
 * it.  dpm3d_bin_width returns the factor (-1 for a bad handle). */
int dpm3d_set_bin_width(dpm_tissue* h, float factor);
float dpm3d_bin_width(dpm_tissue* h);

/* Kernel form of the 3D step (no reference counterpart; the reference
 * launches one work-item per vertex per force term, Tissue3D.cpp:372-445):
 * 1 one workgroup per cell; 2 one 64-lane wave per cell (meshes of at most
 * 256 vertices, several cells per workgroup, no workgroup barriers after
 * set-up); 0 (default) chooses 2 for meshes it takes once the handle holds
 * at least 24 cells per compute unit, else 1.  Both make the same
 * pair decisions; the wave form sums each vertex's edge springs once per ring
 * edge and its adhesion terms in slices, so forces agree to float rounding. */
int dpm3d_set_step_kernel(dpm_tissue* h, int kind);

/* Autotuner of the 3D step (no reference counterpart): times `steps` steps
 * (after one warm-up launch) of each configuration on the handle's current
 * state, restoring that state after every trial, and keeps the fastest.
 * Configurations, in the order of times_ms[8] (ms per step; -1 where a
 * configuration does not apply): workgroup form with the scan, workgroup
 * form with the cell list, then the wave form with 1, 2, 4 cells per
 * workgroup with the scan, and the same with the cell list; each cell-list
 * configuration is timed at bin widths 1, 1.5 and 2 (dpm3d_set_bin_width)
 * and reports its fastest.  best[3] (may be NULL) receives {form, cells per
 * workgroup (0 for the workgroup form), cell list}; the handle keeps them
 * and the winning bin width (dpm3d_bin_width reads it; dpm3d_set_step_kernel,
 * dpm3d_set_cell_list and dpm3d_set_bin_width override).  Results do not
 * depend on the choice beyond float rounding of the grouped sums. */
int dpm3d_autotune(dpm_tissue* h, int steps, float dt, int* best, float* times_ms);

/* Differentiation form (no reference counterpart).  dpm3d_trajectory runs
 * nsteps Euler steps (as dpm_euler) and keeps the positions every `every`
 * steps (nsteps a multiple of every) on the device.  dpm3d_adjoint then
 * computes, by the discrete adjoint of x_{n+1} = x_n + dt F(x_n; theta)
 * with the segments between checkpoints recomputed,
 *   dtheta[c*4 + k] = dL/dtheta_k of cell c, theta = (Kv, Ka, Kb, Ks),
 *   lambda0 = dL/dx_0 (AoS, may be NULL),
 * from lambda_final = dL/dx_N (AoS).  Covers the volume, surface, bending,
 * substrate (with its stiffness gradient) and contact terms; tissues with
 * DPM3D_ADHESION are rejected.  The cut-off decisions of the force law
 * (facing-down faces, substrate heights, contact pairs) are held fixed, as
 * their derivatives vanish almost everywhere.  After the call the positions
 * are those at the end of the trajectory. */
int dpm3d_trajectory(dpm_tissue* h, int nsteps, float dt, int every);
/* Compensated positions for float tissues (no reference counterpart; off
 * by default, where the update is the reference's x += F dt in float): each
 * vertex carries the rounding residual of its position, x + lo, advanced by
 * an exact TwoSum, so that displacements far below one float ulp of the
 * coordinates accumulate instead of being lost.  Forces are evaluated at the
 * float positions x.  dpm3d_get_positions_f64 returns x + lo; the float
 * accessors return x.  Positions set afterwards keep their double part in
 * lo. */
int dpm3d_set_compensated(dpm_tissue* h, int on);
int dpm3d_adjoint(dpm_tissue* h, const float* lambda_final, double* dtheta, float* lambda0);
int dpm3d_adjoint_f64(dpm_tissue* h, const double* lambda_final, double* dtheta, double* lambda0);

/* ------------------------------------------------------------------ 2D ---
 * A batch of tissues of polygon cells with per-cell vertex counts
 * (Tissue2D.cpp:117-140).  Vertices are packed cell after cell (no padding):
 * cell c owns vertices [vert_begin[c], vert_begin[c+1]).
 */
typedef struct {
  int ntissues;
  const int* cell_begin;  /* [ntissues+1] */
  const int* nv;          /* [ncells] vertices per cell (>= 3) */
  /* per cell [ncells] */
  const float* Ka;
  const float* Kl;
  const float* Kb;
  const float* a0;
  const float* l0;
  const float* r0;
  /* per tissue [ntissues] */
  const float* L;
  const int* pbc;
  const float* Kre;
  const float* Kat;
  unsigned mask;          /* DPM2D_* terms (DPM2D_ALL = the reference's) */
} dpm2d_desc;

int dpm2d_create(const dpm2d_desc* desc, dpm_tissue** out);
/* xy[(vert)*2 + k], vertices packed cell after cell, n = total vertices. */
int dpm2d_set_positions(dpm_tissue* h, const float* xy, size_t n);
int dpm2d_get_positions(dpm_tissue* h, float* xy, size_t n);
int dpm2d_get_forces(dpm_tissue* h, float* xy, size_t n);
int dpm2d_set_cell_params(dpm_tissue* h, const float* Ka, const float* Kl,
                          const float* Kb, const float* a0, const float* l0,
                          const float* r0);
int dpm2d_set_tissue_params(dpm_tissue* h, const float* L, const int* pbc,
                            const float* Kre, const float* Kat, unsigned mask);
/* Diagnostic as dpm3d_probe: forces at the current positions; overlap[v] = 1
 * where the repulsion pair exists (Cell2D_kernel.cl:199-205); attraction[v]
 * = number of attraction partners (Cell2D_kernel.cl:261).  On a double
 * tissue only the forces are available (overlap, attraction must be NULL). */
int dpm2d_probe(dpm_tissue* h, float* forces, int* overlap, int* attraction);

/* Diagnostic: per vertex, the wrapping 32-bit sum of the keys f(x ^ f(y))
 * of its attraction partners (Cell2D_kernel.cl:250-266, d < l0 of its cell),
 * the partner identified by the bit patterns of its coordinates (f as in
 * dpm3d_probe_adhesion).  Float tissues. */
int dpm2d_probe_attraction(dpm_tissue* h, unsigned* keys);

/* Step form of the 2D tissues (no reference counterpart; the reference
 * launches its 2D kernels over an NDRange of NCELLS x maxNV work-items,
 * Tissue2D.cpp:212): 1 one workgroup per tissue, the whole call in one launch
 * with the tissue in LDS; 2 the wide form, three launches per step over the
 * whole device on global memory (any tissue size, float tissues); 0
 * (default) the second for tissues beyond one workgroup and for handles of
 * at most 64 tissues whose largest has 128 cells or more, else the first.
 * Both make the same decisions and forces bit for bit.  FIRE,
 * jamming, energies and double tissues need tissues that fit one
 * workgroup. */
int dpm2d_set_step_kernel(dpm_tissue* h, int kind);

/* Double-precision 2D tissue (positions, forces, velocities, box in double;
 * the same formulas), as dpm3d_create_f64.  Its Euler steps, FIRE and
 * jamming all run in double. */
int dpm2d_create_f64(const dpm2d_desc* desc, dpm_tissue** out);
int dpm2d_set_positions_f64(dpm_tissue* h, const double* xy, size_t n);
int dpm2d_get_positions_f64(dpm_tissue* h, double* xy, size_t n);
int dpm2d_get_forces_f64(dpm_tissue* h, double* xy, size_t n);
/* Box length of every tissue (jamming changes it). */
int dpm2d_get_box(dpm_tissue* h, double* L);

/* FIRE minimisation (Bitzek, Koskinen, Gaehler, Moseler & Gumbsch, PRL 97,
 * 170201 (2006)) of every tissue with the tissue's force terms (use
 * DPM2D_CONSERVATIVE: FIRE needs forces that are minus an energy gradient),
 * semi-implicit Euler MD with unit masses, power P = F.v and the norms |F|,
 * |v| reduced in double per tissue; converged when max_v |F_v| < ftol
 * (ftol = 0 runs exactly max_iter iterations).  No
 * reference counterpart (the reference has only the Euler step).  One
 * launch for the whole batch, one workgroup per tissue. */
typedef struct {
  double dt0, dt_max, ftol, f_inc, f_dec, alpha0, f_alpha;
  int max_iter, n_min;
} dpm_fire_params;
/* dt0 0.01, dt_max 0.1, ftol 1e-10, f_inc 1.1, f_dec 0.5, alpha0 0.1,
 * f_alpha 0.99, max_iter 100000, n_min 5. */
void dpm_fire_defaults(dpm_fire_params* p);
/* Per-tissue results of dpm2d_fire / dpm2d_jam / dpm2d_energy. */
typedef struct {
  double iterations;   /* Euler steps or FIRE iterations (all stages) */
  double converged;    /* 1 if the final max |F_v| < ftol */
  double fmax;         /* final max_v |F_v| */
  double energy;       /* final conservative energy (area, perimeter,
                          bending, vertex repulsion) */
  double e_int;        /* its vertex-repulsion part */
  double phi;          /* final packing fraction sum(a0)/L^2 */
  double phi_lo;       /* jamming: largest unjammed phi found */
  double phi_hi;       /* jamming: smallest jammed phi found (0 = none) */
  double stages;       /* jamming: minimisations run */
  double L;            /* final box length */
  double unconverged;  /* minimisations that stopped at max_iter */
  double status;       /* jamming: 0 bracket within phi_tol, 1 phi0 was
                          already jammed, 2 in progress, 3 stopped at
                          max_stages */
  double phi_target;   /* jamming: phi of the next minimisation */
} dpm_relax_stats;
int dpm2d_fire(dpm_tissue* h, const dpm_fire_params* p, dpm_relax_stats* stats);
/* Energy, e_int, fmax, phi, L of the current state (forces updated). */
int dpm2d_energy(dpm_tissue* h, dpm_relax_stats* stats);

/* Jamming by compression and bisection on phi = sum(a0)/L^2 (O'Hern,
 * Silbert, Liu & Nagel, PRE 68, 011306 (2003)): cells are moved rigidly so
 * that their centres scale with the box; at each phi a FIRE minimisation
 * that stops early ("cleared") once the vertex-repulsion energy per cell is
 * at most u_jam; the state is jammed iff it was not cleared (FIRE reached
 * ftol or max_iter with that energy above u_jam).  phi grows by dphi from phi0 until the first jammed state, then
 * bisects between the largest unjammed and smallest jammed phi (restarting
 * from the saved unjammed state) until phi_hi - phi_lo < phi_tol or
 * max_stages minimisations.  Ends on the largest unjammed state.  The whole
 * protocol of every tissue is one launch, captured in a hipGraph that later
 * calls with the same parameters replay.  Needs DPM2D_VREPULSION. */
typedef struct {
  double phi0, dphi, phi_tol, u_jam;
  int max_stages;
  dpm_fire_params fire;
} dpm_jam_params;
/* phi0 0.35, dphi 0.02, phi_tol 1e-3, u_jam 1e-16, max_stages 40, FIRE
 * defaults with max_iter 20000. */
void dpm_jam_defaults(dpm_jam_params* p);
int dpm2d_jam(dpm_tissue* h, const dpm_jam_params* p, dpm_relax_stats* stats);
/* The same protocol over several launches: start != 0 begins it, each call
 * runs at most `stages` more minimisations per tissue (status 2 while one
 * is unfinished; the state stays on the device).  Any split gives results
 * bit-identical to dpm2d_jam; callers use it to report progress or bound a
 * launch's duration. */
int dpm2d_jam_step(dpm_tissue* h, const dpm_jam_params* p, int start, int stages,
                   dpm_relax_stats* stats);
/* Workgroup size of the FIRE / jamming / double-Euler kernel (64..1024, a
 * multiple of 64; 0 = default), and the autotuner that picks it by timing
 * `iters` FIRE iterations of the batch at each size (positions restored);
 * times_ms[5] (may be NULL) receives the timings at 64..1024. */
int dpm2d_set_relax_threads(dpm_tissue* h, int threads);
int dpm2d_relax_threads(dpm_tissue* h);
int dpm2d_autotune(dpm_tissue* h, int iters, int* best, float* times_ms);

/* ---------------------------------------------------------- dispersal ---
 * The soft-disk relaxation of Tissue2D::Disperse and Tissue3D::Disperse2D
 * (Tissue2D.cpp:42-105, Tissue3D.cpp:49-106) on the device, bit-identical to
 * the reference's host loop: n centres (X, Y: in the starting centres, which
 * the reference draws from drand48; out the relaxed ones) with contact radii
 * r (the reference's radius times its scale, 1 in 2D and 2 in 3D) in a
 * periodic square box of side L; pairs closer than r_i + r_j push apart,
 * x += 0.01 F per iteration, until the float energy changes by at most 1e-6
 * or after max_count + 1 iterations (the reference: max_count = 100000).
 * result (may be NULL): [0] iterations run, [1] 1 if stopped by max_count.
 * One workgroup, candidates from a periodic bin grid; n <= 65535 and the
 * centres must fit one workgroup's LDS (about 4,000 centres). */
int dpm_disperse(const float* r, int n, float L, float* X, float* Y, int max_count, int* result);

/* ------------------------------------------------------------- common ---
 * Overdamped forward Euler, x += F*dt, nsteps times (the step loops of
 * Tissue2D.cpp:215-229 and Tissue3D.cpp:372-445).  Validates nsteps > 0 and,
 * for 3D, 0 < dt <= 0.1 (Tissue3D.cpp:123-131).  Synchronous: returns after
 * the device finished; for a 3D tissue the positions are then checked for
 * non-finite values (Tissue3D.cpp:481-488; the 2D reference checks nothing,
 * Tissue2D.cpp:231-240). */
int dpm_euler(dpm_tissue* h, int nsteps, float dt);
/* Same, but asynchronous on the handle's stream with no final check; the
 * caller synchronises (dpm_synchronize) before reading results. */
int dpm_euler_async(dpm_tissue* h, int nsteps, float dt);
int dpm_synchronize(dpm_tissue* h);
/* The handle's HIP stream (hipStream_t as void*), for callers that time the
 * launches with their own events. */
void* dpm_stream(dpm_tissue* h);
/* Counters accumulated since creation or the last reset: [0] steps,
 * [1] (vertex, cell) winding-number evaluations, [2] contact pairs found,
 * [3] candidate cell pairs. */
int dpm_counters(dpm_tissue* h, unsigned long long* out4, int reset);
void dpm_destroy(dpm_tissue* h);

#ifdef __cplusplus
}
#endif
#endif /* DPM_H_ */
