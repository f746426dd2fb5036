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
#define DPM2D_ALL 31u

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
  unsigned mask;          /* DPM2D_* terms */
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
 * = number of attraction partners (Cell2D_kernel.cl:261). */
int dpm2d_probe(dpm_tissue* h, float* forces, int* overlap, int* attraction);

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
