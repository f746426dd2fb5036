// C ABI of the DPM timestep (include/dpm.h): handle management, device
// state, launch orchestration.  One translation unit holds the kernels and
// their host side so that the library is one code object for gfx950.
#include "dpm.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "dpm2d.hip"
#include "dpm3d.hip"
#include "dpm3d_wave.hip"
#include "dpm2d_wide.hip"
#include "dpm_disperse.hip"
#include "dpm_measures.hip"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHECK(call)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(DPM_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_));      \
  } while (0)

template <class T>
int dev_alloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T));
  if (e != hipSuccess) return fail(DPM_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return DPM_OK;
}

template <class T>
int upload(T* dst, const T* src, size_t n, hipStream_t s) {
  HIPCHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
  return DPM_OK;
}

#define TRY(x)                \
  do {                        \
    int rc_ = (x);            \
    if (rc_ != DPM_OK) return rc_; \
  } while (0)

int max_lds_per_block(int dev) {
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    return 65536;
  return v;
}

// Steps replayed per hipGraph launch (even, so the position ping-pong ends
// where it started).
constexpr int kGraphSteps = 64;
// reals per cell of the cached cell statistics (CellStat)
constexpr int kStatWords = (int)(sizeof(CellStat<float>) / sizeof(float));
static_assert(sizeof(CellStat<double>) == kStatWords * sizeof(double), "CellStat layout");

}  // namespace

struct dpm_tissue {
  int dim = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  unsigned long long* counters = nullptr;  // device [4 * DPM_COUNTER_SLOTS]
  std::vector<void*> allocs;

  // ------------------------------------------------------------------ 3D
  int T = 0, C = 0, nv = 0, nf = 0, NT = 0, ncmax = 0;
  int rs = 4;  // bytes per real of positions, forces and cell statistics: 4 or 8
  size_t lds = 0;
  // arrays of the real type (float or double, per rs)
  void *x[2] = {}, *y[2] = {}, *z[2] = {}, *fx = nullptr, *fy = nullptr, *fz = nullptr;
  void* st[2] = {};  // CellStat: centre of mass, bounding radius
  double* vol[2] = {};
  void* em[2] = {};  // per-cell longest edge
  double* energy = nullptr;
  float* meas = nullptr;  // per cell: the host reductions (dpm3d_cell_measures)
  float *Kv = nullptr, *Ka = nullptr, *Ks = nullptr, *Kb = nullptr, *v0 = nullptr, *a0 = nullptr;
  int *cell_tissue = nullptr, *tissue_begin = nullptr;
  float *tL = nullptr, *tKre = nullptr, *tKat = nullptr, *tksg = nullptr;
  int* tpbc = nullptr;
  int *faces = nullptr, *vf = nullptr, *adj = nullptr;
  int ww = 1;
  int closed = 0;
  // wave-per-cell step (meshes of <= 256 vertices): cells per workgroup
  // (0: not available), LDS; step_kernel 0 automatic, 1 workgroup, 2 wave
  int wave_cpw = 0, step_kernel = 0, wave_min_cells = 0;
  void* wtopo = nullptr;  // LDS image of the wave form's shared topology
  int wtopo_n16 = 0;
  size_t lds_w = 0;
  unsigned mask = 0;
  int cur = 0;
  bool stats_valid = false;
  bool forces_valid = false;
  std::vector<int> tissue_begin_h;
  std::vector<int> faces_h;
  hipGraphExec_t graph = nullptr;
  int graph_cur = -1;
  float graph_dt = 0.0f;
  int* flag = nullptr;  // device non-finite flag
  // recorded trajectory for the adjoint
  int traj_steps = 0, traj_every = 0, traj_slots = 3;
  float traj_dt = 0.0f;
  size_t lds_vjp = 0;
  void* ck = nullptr;    // checkpoints [(steps/every + 1) * 3 * NVT] reals
  void* seg = nullptr;   // recomputed segment positions [every * 3 * NVT]
  void* segw = nullptr;  // contact weights of the segment [every * NVT]
  void* lam = nullptr;   // adjoint [3 * NVT]
  double* gtheta = nullptr;  // [C * 4]
  void* segwp = nullptr;     // recorded partner winding numbers [every * NVT * MAXPART]
  void* segsh = nullptr;     // recorded partner shifts [every * NVT * MAXPART * 3]
  int* segcell = nullptr;    // recorded partner cells [every * NVT * MAXPART]
  void* cross = nullptr;     // other cells' adjoint contributions [3 * NVT]
  int* rec_overflow = nullptr;
  size_t ck_cap = 0, seg_cap = 0, segw_cap = 0, lam_cap = 0, segwp_cap = 0, segsh_cap = 0, cross_cap = 0,
         segcell_cap = 0;
  void* lo = nullptr;  // compensated positions: residuals [3 * NVT] (nullptr: off)
  // cell list (counting sort of the centres of mass into hashed bins)
  int cell_list_mode = -1;  // -1 automatic, 0 off, 1 on
  bool cells_on = false;
  float bin_scale = 1.0f;  // cell-list bin width over the step's largest candidate reach
  int *bin_start = nullptr, *bin_list = nullptr, *tbin_nb = nullptr, *tbin_nbuck = nullptr,
      *tbin_hoff = nullptr;
  size_t bin_lds = 0;

  // ------------------------------------------------------------------ 2D
  int V = 0, NT2 = 0, NTr = 0;  // Euler threads, relax threads per tissue
  size_t lds2 = 0;
  size_t lds_run = 0;  // k2d_run: lds2 and, where it fits, the crossing-parity pair list
  int over_cap = 0;
  // arrays of the real type (rs bytes)
  void *x2 = nullptr, *y2 = nullptr, *fx2 = nullptr, *fy2 = nullptr;
  void *vx2 = nullptr, *vy2 = nullptr, *sx2 = nullptr, *sy2 = nullptr;  // FIRE / jamming
  void* tLr = nullptr;  // per-tissue box length
  double* res2 = nullptr;     // [T * RS_COUNT]
  double* sum_a0 = nullptr;   // [T]
  int *tcb = nullptr, *cvb = nullptr;
  float *Kl = nullptr, *l0_2 = nullptr, *r0 = nullptr;
  std::vector<int> cvb_h;
  std::vector<float> a0_h;
  hipGraphExec_t jam_graph = nullptr;
  std::vector<unsigned char> jam_key;
  // tissues beyond one workgroup: 0 automatic, 1 one workgroup per tissue,
  // 2 the wide path (three or four launches per step over the device)
  int wide_mode = 0;
  bool fits_wg = true;  // every tissue fits one workgroup (k2d_run)
  int max_cells = 0;    // cells of the largest tissue
  float *xw = nullptr, *yw = nullptr, *cellq = nullptr;
  int *ncand_w = nullptr, *cell_tissue2 = nullptr, *vbl = nullptr;
  short* cand_w = nullptr;
  unsigned* over_w = nullptr;  // crossing-parity bits of the wide path
  int* vcell_w = nullptr;      // cell of each vertex
  hipGraphExec_t wide_graph = nullptr;
  float wide_graph_dt = 0.0f;

  template <class T_>
  int alloc(T_** p, size_t n) {
    TRY(dev_alloc(p, n));
    allocs.push_back(*p);
    return DPM_OK;
  }
  int alloc_real(void** p, size_t n) {  // n reals of rs bytes
    unsigned char* q = nullptr;
    TRY(alloc(&q, n * (size_t)rs));
    *p = q;
    return DPM_OK;
  }
  ~dpm_tissue() {
    if (graph) (void)hipGraphExecDestroy(graph);
    if (jam_graph) (void)hipGraphExecDestroy(jam_graph);
    if (wide_graph) (void)hipGraphExecDestroy(wide_graph);
    for (void* p : allocs) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// ------------------------------------------------------------------ common

extern "C" const char* dpm_last_error(void) { return g_err.c_str(); }
extern "C" int dpm_version(void) { return 1; }

extern "C" int dpm_device_count(int* n) {
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  if (n) *n = c;
  return DPM_OK;
}

static int init_common(dpm_tissue* h) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    return fail(DPM_EDEVICE, "no HIP device is visible");
  HIPCHECK(hipGetDevice(&h->device));
  HIPCHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  TRY(h->alloc(&h->counters, 4 * DPM_COUNTER_SLOTS));
  TRY(h->alloc(&h->flag, 1));
  HIPCHECK(hipMemsetAsync(h->counters, 0, 4 * DPM_COUNTER_SLOTS * sizeof(unsigned long long), h->stream));
  return DPM_OK;
}

extern "C" void dpm_destroy(dpm_tissue* h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  delete h;
}

extern "C" void* dpm_stream(dpm_tissue* h) { return h ? (void*)h->stream : nullptr; }

extern "C" int dpm_synchronize(dpm_tissue* h) {
  if (!h) return fail(DPM_EINVAL, "null handle");
  HIPCHECK(hipStreamSynchronize(h->stream));
  return DPM_OK;
}

extern "C" int dpm_counters(dpm_tissue* h, unsigned long long* out4, int reset) {
  if (!h || !out4) return fail(DPM_EINVAL, "null argument");
  std::vector<unsigned long long> slots(4 * DPM_COUNTER_SLOTS);
  HIPCHECK(hipMemcpyAsync(slots.data(), h->counters, slots.size() * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  for (int k = 0; k < 4; k++) {
    out4[k] = 0;
    for (int s = 0; s < DPM_COUNTER_SLOTS; s++) out4[k] += slots[4 * s + k];
  }
  if (reset)
    HIPCHECK(hipMemsetAsync(h->counters, 0, slots.size() * sizeof(unsigned long long), h->stream));
  return DPM_OK;
}

template <typename R>
__global__ void k_nonfinite(const R* a, const R* b, const R* c, size_t n, int* flag) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const bool bad = !isfinite(a[i]) || !isfinite(b[i]) || (c && !isfinite(c[i]));
    if (bad) atomicOr(flag, 1);
  }
}

// Positions of all vertices finite?  (Tissue3D.cpp:481-488)
template <typename R>
static int check_finite(dpm_tissue* h, const R* a, const R* b, const R* c, size_t n) {
  HIPCHECK(hipMemsetAsync(h->flag, 0, sizeof(int), h->stream));
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 2048);
  k_nonfinite<R><<<blocks, 256, 0, h->stream>>>(a, b, c, n, h->flag);
  HIPCHECK(hipGetLastError());
  int f = 0;
  HIPCHECK(hipMemcpyAsync(&f, h->flag, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  if (f) return fail(DPM_ENONFINITE, "Non-finite simulation results");
  return DPM_OK;
}

// ================================================================== 3D ====

template <typename R>
static bool all_finite(const R* p, size_t n) {
  for (size_t i = 0; i < n; i++)
    if (!std::isfinite(p[i])) return false;
  return true;
}

// The wave-per-cell step kernel of a precision.
// The wave form's kernel for a handle: the compile-time-sized instance for
// float 162-vertex, 320-face meshes, else the general one.
static const void* wstep_kernel(int rs, int nv, int nf, bool probe = false) {
  if (rs == 4 && nv == 162 && nf == 320) return probe ? (const void*)k3d_wstep_f32_v162_probe
                                                      : (const void*)k3d_wstep_f32_v162;
  return rs == 8 ? (const void*)k3d_wstep_f64 : (const void*)k3d_wstep_f32;
}

static int create3d(const dpm3d_desc* d, dpm_tissue** out, int rs) {
  if (!d || !out) return fail(DPM_EINVAL, "null argument");
  *out = nullptr;
  if (d->ntissues <= 0 || !d->cell_begin) return fail(DPM_EINVAL, "NCELLS must be positive");
  const int T = d->ntissues;
  if (d->cell_begin[0] != 0) return fail(DPM_EINVAL, "cell_begin[0] must be 0");
  int ncmax = 0;
  for (int t = 0; t < T; t++) {
    const int n = d->cell_begin[t + 1] - d->cell_begin[t];
    if (n <= 0) return fail(DPM_EINVAL, "NCELLS must be positive");
    ncmax = std::max(ncmax, n);
  }
  const int C = d->cell_begin[T];
  const int nv = d->nv, nf = d->nf;
  if (nv < 4 || nf < 4 || !d->faces) return fail(DPM_EINVAL, "mesh needs nv >= 4 and nf >= 4");
  if (nv > 1024) return fail(DPM_ERANGE, "nv > 1024 vertices per cell is not supported");
  if (!d->Kv || !d->Ka || !d->Ks || !d->v0 || !d->a0 || !d->L || !d->pbc || !d->Kre || !d->Kat)
    return fail(DPM_EINVAL, "missing parameter array");
  for (int f = 0; f < nf; f++)
    for (int k = 0; k < 3; k++) {
      const int v = d->faces[3 * f + k];
      if (v < 0 || v >= nv) {
        char buf[160];
        snprintf(buf, sizeof(buf), "Invalid face indices (face %d: %d,%d,%d)", f, d->faces[3 * f],
                 d->faces[3 * f + 1], d->faces[3 * f + 2]);
        return fail(DPM_ERANGE, buf);
      }
    }
  for (int c = 0; c < C; c++) {
    if (!(d->Kv[c] > 0) || !(d->Ka[c] > 0) || !(d->Ks[c] > 0))
      return fail(DPM_ERANGE, "Invalid spring constants");
    if (!(d->v0[c] > 0) || !(d->a0[c] > 0)) return fail(DPM_ERANGE, "Invalid reference values");
  }
  // closed two-manifold mesh: every undirected edge on exactly two faces
  std::vector<int> ecount((size_t)nv * nv, 0);
  for (int f = 0; f < nf; f++)
    for (int k = 0; k < 3; k++) {
      const int u = d->faces[3 * f + k], w = d->faces[3 * f + (k + 1) % 3];
      ecount[(size_t)std::min(u, w) * nv + std::max(u, w)]++;
    }
  bool closed = true;
  for (int c : ecount) closed = closed && (c == 0 || c == 2);
  // incidence (face*3 + corner) and adjacency, ascending
  std::vector<std::vector<int>> inc(nv), nb(nv);
  for (int f = 0; f < nf; f++)
    for (int k = 0; k < 3; k++) {
      const int v = d->faces[3 * f + k];
      inc[v].push_back(3 * f + k);
      const int a = d->faces[3 * f + (k + 1) % 3];
      nb[v].push_back(a);
      nb[a].push_back(v);
    }
  std::vector<int> vf((size_t)nv * DPM3D_MAXDEG, -1), adj((size_t)nv * DPM3D_MAXDEG, -1);
  for (int v = 0; v < nv; v++) {
    std::sort(nb[v].begin(), nb[v].end());
    nb[v].erase(std::unique(nb[v].begin(), nb[v].end()), nb[v].end());
    if ((int)inc[v].size() > DPM3D_MAXDEG || (int)nb[v].size() > DPM3D_MAXDEG)
      return fail(DPM_ERANGE, "vertex degree above 8 is not supported");
    for (size_t k = 0; k < inc[v].size(); k++) vf[(size_t)v * DPM3D_MAXDEG + k] = inc[v][k];
    for (size_t k = 0; k < nb[v].size(); k++) adj[(size_t)v * DPM3D_MAXDEG + k] = nb[v][k];
  }
  // Neighbour rows list each vertex's ring u_0 .. u_(k-1) so that face q is
  // (v, u_q, u_(q+1)) in its corner order (the wave form walks it once per
  // vertex; every form sums bending in this order).  That needs every
  // vertex's faces to form one cycle; otherwise the mesh takes the
  // per-corner path (closed = 0) and the rows stay ascending.
  std::vector<int> ring = adj;
  for (int v = 0; v < nv && closed; v++) {
    const size_t k = inc[v].size();
    if (k < 3 || nb[v].size() != k) {
      closed = false;
      break;
    }
    std::vector<std::pair<int, int>> ab;  // (next corner, the one after) per face
    for (int e : inc[v]) {
      const int f = e / 3, c = e % 3;
      ab.push_back({d->faces[3 * f + (c + 1) % 3], d->faces[3 * f + (c + 2) % 3]});
    }
    int u = ab[0].first;
    std::vector<int> r;
    for (size_t q = 0; q < k && closed; q++) {
      r.push_back(u);
      int b = -1, hits = 0;
      for (const auto& x : ab)
        if (x.first == u) {
          b = x.second;
          hits++;
        }
      closed = hits == 1;
      u = b;
    }
    std::vector<int> rs_sorted = r;
    std::sort(rs_sorted.begin(), rs_sorted.end());
    closed = closed && u == ab[0].first && rs_sorted == nb[v];
    for (size_t q = 0; q < k && closed; q++) ring[(size_t)v * DPM3D_MAXDEG + q] = r[q];
  }
  if (closed) adj = ring;

  dpm_tissue* h = new dpm_tissue();
  h->dim = 3;
  h->rs = rs;
  int rc = init_common(h);
  if (rc) { delete h; return rc; }
  h->T = T;
  h->C = C;
  h->nv = nv;
  h->nf = nf;
  h->closed = closed ? 1 : 0;
  h->ncmax = ncmax;
  h->NT = std::max(256, (nv + 63) / 64 * 64);
  // Waves that evaluate winding numbers, each with an nv-entry direction
  // table in LDS (at most ~48 KB of tables).
  h->ww = std::max(1, std::min(h->NT / 64, 49152 / (4 * rs * nv)));
  h->lds = lds3d_layout(nv, nf, h->NT, h->ww, rs).total;
  if ((int)h->lds > max_lds_per_block(h->device)) {
    delete h;
    return fail(DPM_ERANGE, "cell mesh does not fit the LDS of one workgroup");
  }
  const size_t NVT = (size_t)C * nv;
  auto A = [&](auto** p, size_t n) { return h->alloc(p, n); };
  rc = 0;
  auto AR = [&](void** p, size_t n) { return h->alloc_real(p, n); };
  for (int b = 0; b < 2 && !rc; b++) {
    rc |= AR(&h->x[b], NVT);
    rc |= AR(&h->y[b], NVT);
    rc |= AR(&h->z[b], NVT);
    rc |= AR(&h->st[b], kStatWords * (size_t)C);
    rc |= A(&h->vol[b], (size_t)C);
    rc |= AR(&h->em[b], (size_t)C);
  }
  rc |= AR(&h->fx, NVT);
  rc |= AR(&h->fy, NVT);
  rc |= AR(&h->fz, NVT);
  rc |= A(&h->energy, (size_t)C);
  rc |= A(&h->Kv, C);
  rc |= A(&h->Ka, C);
  rc |= A(&h->Ks, C);
  rc |= A(&h->Kb, C);
  rc |= A(&h->v0, C);
  rc |= A(&h->a0, C);
  rc |= A(&h->cell_tissue, C);
  rc |= A(&h->tissue_begin, T + 1);
  rc |= A(&h->tL, T);
  rc |= A(&h->tpbc, T);
  rc |= A(&h->tKre, T);
  rc |= A(&h->tKat, T);
  rc |= A(&h->tksg, T);
  rc |= A(&h->faces, (size_t)nf * 3);
  rc |= A(&h->vf, vf.size());
  rc |= A(&h->adj, adj.size());
  // cell-list buckets per tissue: a power of two >= 2 cells (<= 8192; larger
  // tissues scan every cell)
  std::vector<int> nbuck(T), hoff(T);
  size_t nbtot = 0, maxb = 0;
  int ncmax_t = 0;
  for (int t = 0; t < T; t++) {
    const int nc = d->cell_begin[t + 1] - d->cell_begin[t];
    ncmax_t = std::max(ncmax_t, nc);
    int nbk = 64;
    while (nbk < 2 * nc) nbk *= 2;
    if (nbk > 8192) nbk = 0;
    nbuck[t] = nbk;
    hoff[t] = (int)nbtot;
    nbtot += (size_t)nbk + 1;
    maxb = std::max(maxb, (((size_t)nbk + 1 + 2 * (size_t)nc) * 4 + 7) / 8 * 8);
  }
  h->bin_lds = maxb + 32 * 8;
  rc |= A(&h->bin_start, nbtot);
  rc |= A(&h->bin_list, (size_t)C);
  rc |= A(&h->tbin_nb, (size_t)T);
  rc |= A(&h->tbin_nbuck, (size_t)T);
  rc |= A(&h->tbin_hoff, (size_t)T);
  if (rc) { delete h; return DPM_ENOMEM; }
  h->tissue_begin_h.assign(d->cell_begin, d->cell_begin + T + 1);
  h->faces_h.assign(d->faces, d->faces + 3 * nf);
  std::vector<int> ct(C);
  for (int t = 0; t < T; t++)
    for (int c = d->cell_begin[t]; c < d->cell_begin[t + 1]; c++) ct[c] = t;
  hipStream_t s = h->stream;
  rc = 0;
  rc |= upload(h->cell_tissue, ct.data(), C, s);
  rc |= upload(h->tissue_begin, d->cell_begin, T + 1, s);
  rc |= upload(h->faces, d->faces, (size_t)nf * 3, s);
  rc |= upload(h->vf, vf.data(), vf.size(), s);
  rc |= upload(h->adj, adj.data(), adj.size(), s);
  rc |= upload(h->tbin_nbuck, nbuck.data(), (size_t)T, s);
  rc |= upload(h->tbin_hoff, hoff.data(), (size_t)T, s);
  // automatic: the cell list pays only where a cell would scan many chunks
  // (at 4096 cells x 642 vertices it measured 2.57 ms against the scan's
  // 2.43 ms per step)
  h->cells_on = ncmax_t > 8 * h->NT;
  HIPCHECK(hipMemsetAsync(h->fx, 0, NVT * rs, s));
  HIPCHECK(hipMemsetAsync(h->fy, 0, NVT * rs, s));
  HIPCHECK(hipMemsetAsync(h->fz, 0, NVT * rs, s));
  if (rc) { delete h; return rc; }
  rc = dpm3d_set_cell_params(h, d->Kv, d->Ka, d->Ks, d->Kb, d->v0, d->a0);
  std::vector<float> no_grad(T, 0.0f);
  if (!rc)
    rc = dpm3d_set_tissue_params(h, d->L, d->pbc, d->Kre, d->Kat, d->ks_grad ? d->ks_grad : no_grad.data(),
                                 d->mask);
  if (!d->Kb && !rc) {
    std::vector<float> z(C, 0.0f);
    rc = upload(h->Kb, z.data(), C, s);
  }
  if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = fail(DPM_EDEVICE, "stream sync failed");
  if (rc) { delete h; return rc; }
  const void* kern[2][3] = {{(const void*)k3d_step_f32, (const void*)k3d_stats_f32, (const void*)k3d_energy_f32},
                            {(const void*)k3d_step_f64, (const void*)k3d_stats_f64, (const void*)k3d_energy_f64}};
  for (const void* k : kern[rs == 8 ? 1 : 0])
    (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->lds);
  (void)hipFuncSetAttribute((const void*)k3d_step_f32_v162, hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->lds);
  (void)hipFuncSetAttribute((const void*)k3d_step_f32_v162_probe, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)h->lds);
  (void)hipFuncSetAttribute((const void*)k3d_step_f32_v642, hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->lds);
  (void)hipFuncSetAttribute((const void*)k3d_step_f64_v642, hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->lds);
  (void)hipFuncSetAttribute((const void*)k3d_step_f64_v162, hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->lds);
  (void)hipFuncSetAttribute(rs == 8 ? (const void*)k3d_bin_f64 : (const void*)k3d_bin_f32,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->bin_lds);
  if (nv <= kWaveMaxNv) {  // cells per workgroup: as many as fit 64 KiB of LDS, at most 4
    int cpw = 4;
    while (cpw > 1 && ldsw_layout(nv, nf, cpw, rs).total > 65536) cpw /= 2;
    if (ldsw_layout(nv, nf, cpw, rs).total <= 65536) {
      h->wave_cpw = cpw;
      h->lds_w = ldsw_layout(nv, nf, cpw, rs).total;
      // automatic choice: the wave form once the batch gives every CU about
      // 24 cells (measured crossover between 16 and 32 cells per CU on the
      // 256-cell monolayer); below, the workgroup form fills the device better
      int dev = 0, cus = 256;
      if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
      h->wave_min_cells = 24 * std::max(1, cus);
      // the topology image: faces packed 10 bits per index (staged in
      // LDS), incidence and neighbour rows of 16-bit entries (0xffff: none)
      const Lds3DWave LW = ldsw_layout(nv, nf, cpw, rs);
      std::vector<unsigned char> img(LW.img, 0);
      unsigned* pf = reinterpret_cast<unsigned*>(img.data());
      for (int f = 0; f < nf; f++)
        pf[f] = (unsigned)d->faces[3 * f] | ((unsigned)d->faces[3 * f + 1] << 10) |
                ((unsigned)d->faces[3 * f + 2] << 20);
      unsigned short* pv = reinterpret_cast<unsigned short*>(img.data() + LW.vf);
      unsigned short* pa = reinterpret_cast<unsigned short*>(img.data() + LW.adj);
      for (size_t i = 0; i < vf.size(); i++) {
        pv[i] = (unsigned short)vf[i];
        pa[i] = (unsigned short)adj[i];
      }
      h->wtopo_n16 = (int)(img.size() / 16);
      unsigned char* q = nullptr;
      if (h->alloc(&q, img.size()) || hipMemcpy(q, img.data(), img.size(), hipMemcpyHostToDevice) != hipSuccess) {
        delete h;
        return fail(DPM_EDEVICE, "topology upload failed");
      }
      h->wtopo = q;
      for (bool probe : {false, true})
        (void)hipFuncSetAttribute(wstep_kernel(rs, nv, nf, probe), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)h->lds_w);
    }
  }
  h->lds_vjp = ldsvjp_layout(nv, nf, rs).total;
  (void)hipFuncSetAttribute(rs == 8 ? (const void*)k3d_vjp_f64 : (const void*)k3d_vjp_f32,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->lds_vjp);
  if (rs == 8)
    (void)hipFuncSetAttribute((const void*)k3d_vjp_f64_t256, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)h->lds_vjp);
  *out = h;
  return DPM_OK;
}

extern "C" int dpm3d_set_bin_width(dpm_tissue* h, float factor) {
  if (!h || h->dim != 3) return fail(DPM_EINVAL, "not a 3D tissue");
  if (!(factor >= 1.0f && factor <= 4.0f)) return fail(DPM_EINVAL, "the bin width factor must be in [1, 4]");
  h->bin_scale = factor;
  if (h->graph) {  // captured with the old arguments
    (void)hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  return DPM_OK;
}

extern "C" float dpm3d_bin_width(dpm_tissue* h) { return h && h->dim == 3 ? h->bin_scale : -1.0f; }

extern "C" int dpm3d_set_cell_list(dpm_tissue* h, int mode) {
  if (!h || h->dim != 3) return fail(DPM_EINVAL, "not a 3D tissue");
  if (mode < -1 || mode > 1) return fail(DPM_EINVAL, "mode must be -1 (automatic), 0 or 1");
  int ncmax_t = 0;
  for (int t = 0; t < h->T; t++) ncmax_t = std::max(ncmax_t, h->tissue_begin_h[t + 1] - h->tissue_begin_h[t]);
  h->cell_list_mode = mode;
  h->cells_on = mode == 1 || (mode == -1 && ncmax_t > 8 * h->NT);
  if (h->graph) {  // captured with the old arguments
    (void)hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  return DPM_OK;
}

extern "C" int dpm3d_set_step_kernel(dpm_tissue* h, int kind) {
  if (!h || h->dim != 3) return fail(DPM_EINVAL, "not a 3D tissue");
  if (kind < 0 || kind > 2) return fail(DPM_EINVAL, "kind must be 0 (automatic), 1 (workgroup) or 2 (wave)");
  if (kind == 2 && !h->wave_cpw) return fail(DPM_EINVAL, "the wave-per-cell step needs at most 256 vertices per cell");
  h->step_kernel = kind;
  if (h->graph) {  // captured with the other kernel
    (void)hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  return DPM_OK;
}

extern "C" int dpm3d_create(const dpm3d_desc* d, dpm_tissue** out) { return create3d(d, out, 4); }
extern "C" int dpm3d_create_f64(const dpm3d_desc* d, dpm_tissue** out) { return create3d(d, out, 8); }

extern "C" int dpm3d_set_cell_params(dpm_tissue* h, const float* Kv, const float* Ka,
                                     const float* Ks, const float* Kb, const float* v0,
                                     const float* a0) {
  if (!h || h->dim != 3) return fail(DPM_EINVAL, "not a 3D tissue");
  const int C = h->C;
  for (int c = 0; c < C; c++) {
    if ((Kv && !(Kv[c] > 0)) || (Ka && !(Ka[c] > 0)) || (Ks && !(Ks[c] > 0)))
      return fail(DPM_ERANGE, "Invalid spring constants");
    if ((v0 && !(v0[c] > 0)) || (a0 && !(a0[c] > 0)))
      return fail(DPM_ERANGE, "Invalid reference values");
  }
  hipStream_t s = h->stream;
  if (Kv) TRY(upload(h->Kv, Kv, C, s));
  if (Ka) TRY(upload(h->Ka, Ka, C, s));
  if (Ks) TRY(upload(h->Ks, Ks, C, s));
  if (Kb) TRY(upload(h->Kb, Kb, C, s));
  if (v0) TRY(upload(h->v0, v0, C, s));
  if (a0) {
    TRY(upload(h->a0, a0, C, s));
    h->stats_valid = false;  // the cached statistics carry each cell's l0
  }
  HIPCHECK(hipStreamSynchronize(s));
  return DPM_OK;
}

extern "C" int dpm3d_set_tissue_params(dpm_tissue* h, const float* L, const int* pbc,
                                       const float* Kre, const float* Kat, const float* ks_grad,
                                       unsigned mask) {
  if (!h || h->dim != 3) return fail(DPM_EINVAL, "not a 3D tissue");
  const int T = h->T;
  hipStream_t s = h->stream;
  if (L) {
    for (int t = 0; t < T; t++)
      if (!(L[t] > 0) || !std::isfinite(L[t])) return fail(DPM_EINVAL, "box length must be positive");
    TRY(upload(h->tL, L, T, s));
  }
  if (pbc) TRY(upload(h->tpbc, pbc, T, s));
  if (Kre) TRY(upload(h->tKre, Kre, T, s));
  if (Kat) TRY(upload(h->tKat, Kat, T, s));
  if (ks_grad) TRY(upload(h->tksg, ks_grad, T, s));
  h->mask = mask & DPM3D_ALL;
  HIPCHECK(hipStreamSynchronize(s));
  if (h->graph) {
    (void)hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  return DPM_OK;
}

// Host AoS (reference order) <-> device SoA of the handle's real type.
template <typename In>
static int set_positions3(dpm_tissue* h, const In* xyz, size_t n) {
  if (!h || h->dim != 3 || !xyz) return fail(DPM_EINVAL, "bad arguments");
  const size_t NVT = (size_t)h->C * h->nv;
  if (n != NVT) return fail(DPM_EINVAL, "position count does not match the tissue");
  if (!all_finite(xyz, 3 * n)) return fail(DPM_ENONFINITE, "Non-finite vertex coordinates");
  const int k = h->cur;
  auto put = [&](auto tag) -> int {
    using R = decltype(tag);
    std::vector<R> a(n), b(n), c(n);
    for (size_t i = 0; i < n; i++) {
      a[i] = (R)xyz[3 * i];
      b[i] = (R)xyz[3 * i + 1];
      c[i] = (R)xyz[3 * i + 2];
    }
    TRY(upload((R*)h->x[k], a.data(), n, h->stream));
    TRY(upload((R*)h->y[k], b.data(), n, h->stream));
    TRY(upload((R*)h->z[k], c.data(), n, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
    return DPM_OK;
  };
  TRY(h->rs == 8 ? put(0.0) : put(0.0f));
  if (h->lo) {  // residuals: what the float positions did not keep
    std::vector<float> r(3 * n);
    for (size_t i = 0; i < n; i++)
      for (int k = 0; k < 3; k++)
        r[k * n + i] = (float)((double)xyz[3 * i + k] - (double)(float)xyz[3 * i + k]);
    TRY(upload((float*)h->lo, r.data(), 3 * n, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
  }
  h->stats_valid = false;
  return DPM_OK;
}

template <typename Out>
static int download3(dpm_tissue* h, const void* a, const void* b, const void* c, Out* xyz, size_t n) {
  auto get = [&](auto tag) -> int {
    using R = decltype(tag);
    std::vector<R> A(n), B(n), Cc(n);
    HIPCHECK(hipMemcpyAsync(A.data(), a, n * sizeof(R), hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipMemcpyAsync(B.data(), b, n * sizeof(R), hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipMemcpyAsync(Cc.data(), c, n * sizeof(R), hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
    for (size_t i = 0; i < n; i++) {
      xyz[3 * i] = (Out)A[i];
      xyz[3 * i + 1] = (Out)B[i];
      xyz[3 * i + 2] = (Out)Cc[i];
    }
    return DPM_OK;
  };
  return h->rs == 8 ? get(0.0) : get(0.0f);
}

template <typename Out>
static int get_positions3(dpm_tissue* h, Out* xyz, size_t n) {
  if (!h || h->dim != 3 || !xyz) return fail(DPM_EINVAL, "bad arguments");
  if (n != (size_t)h->C * h->nv) return fail(DPM_EINVAL, "position count does not match the tissue");
  const int k = h->cur;
  TRY(download3(h, h->x[k], h->y[k], h->z[k], xyz, n));
  if (h->lo && sizeof(Out) == 8) {  // compensated: x + residual
    std::vector<Out> r(3 * n);
    TRY(download3(h, h->lo, (float*)h->lo + n, (float*)h->lo + 2 * n, r.data(), n));
    for (size_t i = 0; i < 3 * n; i++) xyz[i] += r[i];
  }
  return DPM_OK;
}

extern "C" int dpm3d_set_compensated(dpm_tissue* h, int on) {
  if (!h || h->dim != 3) return fail(DPM_EINVAL, "not a 3D tissue");
  if (h->rs != 4) return fail(DPM_EINVAL, "compensated positions apply to float tissues");
  const size_t n = 3 * (size_t)h->C * h->nv;
  if (on && !h->lo) {
    float* p = nullptr;
    TRY(h->alloc(&p, n));
    h->lo = p;
    HIPCHECK(hipMemsetAsync(h->lo, 0, n * sizeof(float), h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
  } else if (!on && h->lo) {
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(h->lo);
    h->allocs.erase(std::remove(h->allocs.begin(), h->allocs.end(), h->lo), h->allocs.end());
    h->lo = nullptr;
  }
  if (h->graph) {  // captured with the old arguments
    (void)hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  return DPM_OK;
}

template <typename Out>
static int get_forces3(dpm_tissue* h, Out* xyz, size_t n) {
  if (!h || h->dim != 3 || !xyz) return fail(DPM_EINVAL, "bad arguments");
  if (n != (size_t)h->C * h->nv) return fail(DPM_EINVAL, "force count does not match the tissue");
  return download3(h, h->fx, h->fy, h->fz, xyz, n);
}

extern "C" int dpm3d_set_positions(dpm_tissue* h, const float* xyz, size_t n) {
  return set_positions3(h, xyz, n);
}
extern "C" int dpm3d_get_positions(dpm_tissue* h, float* xyz, size_t n) {
  return get_positions3(h, xyz, n);
}
extern "C" int dpm3d_get_forces(dpm_tissue* h, float* xyz, size_t n) { return get_forces3(h, xyz, n); }
extern "C" int dpm3d_set_positions_f64(dpm_tissue* h, const double* xyz, size_t n) {
  return set_positions3(h, xyz, n);
}
extern "C" int dpm3d_get_positions_f64(dpm_tissue* h, double* xyz, size_t n) {
  return get_positions3(h, xyz, n);
}
extern "C" int dpm3d_get_forces_f64(dpm_tissue* h, double* xyz, size_t n) {
  return get_forces3(h, xyz, n);
}

template <typename R>
static Step3D<R> step_args(dpm_tissue* h, int in, float dt, int flags) {
  Step3D<R> a{};
  a.xin = (const R*)h->x[in];
  a.yin = (const R*)h->y[in];
  a.zin = (const R*)h->z[in];
  a.xout = (R*)h->x[1 - in];
  a.yout = (R*)h->y[1 - in];
  a.zout = (R*)h->z[1 - in];
  a.fx = (R*)h->fx;
  a.fy = (R*)h->fy;
  a.fz = (R*)h->fz;
  a.st_in = (const CellStat<R>*)h->st[in];
  a.vol_in = h->vol[in];
  a.em_in = (const R*)h->em[in];
  a.st_out = (CellStat<R>*)h->st[1 - in];
  a.vol_out = h->vol[1 - in];
  a.em_out = (R*)h->em[1 - in];
  a.Kv = h->Kv;
  a.Ka = h->Ka;
  a.Ks = h->Ks;
  a.Kb = h->Kb;
  a.v0 = h->v0;
  a.a0 = h->a0;
  a.cell_tissue = h->cell_tissue;
  a.tissue_begin = h->tissue_begin;
  a.tL = h->tL;
  a.tpbc = h->tpbc;
  a.tKre = h->tKre;
  a.tKat = h->tKat;
  a.tksg = h->tksg;
  a.faces = h->faces;
  a.vf = h->vf;
  a.adj = h->adj;
  a.ww = h->ww;
  a.nv = h->nv;
  a.nf = h->nf;
  a.ncells = h->C;
  a.closed = h->closed;
  a.wtopo = (const uint4*)h->wtopo;
  a.wtopo_n16 = h->wtopo_n16;
  a.dt = (R)dt;
  a.mask = h->mask;
  a.flags = flags;
  a.counters = h->counters;
  a.ncmax = h->ncmax;
  a.energy = h->energy;
  if (h->lo) {
    a.xlo = (R*)h->lo;
    a.ylo = (R*)h->lo + (size_t)h->C * h->nv;
    a.zlo = (R*)h->lo + 2 * (size_t)h->C * h->nv;
  }
  if (h->cells_on) {
    a.bin_start = h->bin_start;
    a.bin_list = h->bin_list;
    a.tbin_nb = h->tbin_nb;
    a.tbin_nbuck = h->tbin_nbuck;
    a.tbin_hoff = h->tbin_hoff;
  }
  return a;
}

enum Kind3D { K_STEP, K_STATS, K_ENERGY };

// Launches one kernel of the 3D family over all cells, in the handle's
// precision; `edit` may adjust the arguments before the launch.
template <typename Edit>
static int launch3d(dpm_tissue* h, Kind3D kind, int in, float dt, int flags, Edit edit) {
  auto go = [&](auto tag) -> int {
    using R = decltype(tag);
    Step3D<R> a = step_args<R>(h, in, dt, flags);
    edit(a);
    void (*k)(Step3D<R>) = nullptr;
    if constexpr (sizeof(R) == 8)
      k = kind == K_STEP ? k3d_step_f64 : kind == K_STATS ? k3d_stats_f64 : k3d_energy_f64;
    else
      k = kind == K_STEP ? k3d_step_f32 : kind == K_STATS ? k3d_stats_f32 : k3d_energy_f32;
    // diagnostic outputs requested: the instances that record them
    const bool probe = a.p_contact || a.p_winding || a.p_down || a.p_adh || a.p_adh_hash || a.p_contact_key;
    if constexpr (sizeof(R) == 4)
      if (kind == K_STEP && h->ww == 4 && h->nv == 162 && h->nf == 320 && h->NT == 256)
        k = probe ? k3d_step_f32_v162_probe : k3d_step_f32_v162;
    if constexpr (sizeof(R) == 4)
      if (kind == K_STEP && h->ww == 4 && h->nv == 642 && h->nf == 1280 && h->NT == 704) k = k3d_step_f32_v642;
    if constexpr (sizeof(R) == 8)
      if (kind == K_STEP && h->ww == 2 && h->nv == 642 && h->nf == 1280 && h->NT == 704) k = k3d_step_f64_v642;
    if constexpr (sizeof(R) == 8)
      if (kind == K_STEP && h->ww == 4 && h->nv == 162 && h->nf == 320 && h->NT == 256) k = k3d_step_f64_v162;
    if (kind == K_STEP && a.bin_list) {  // this step's cell list, from its input statistics
      Bin3D<R> b{};
      b.tissue_begin = h->tissue_begin;
      b.st_in = a.st_in;
      b.em_in = a.em_in;
      b.a0 = h->a0;
      b.tL = h->tL;
      b.tpbc = h->tpbc;
      b.tKre = h->tKre;
      b.tKat = h->tKat;
      b.mask = h->mask;
      b.tbin_nb = h->tbin_nb;
      b.tbin_nbuck = h->tbin_nbuck;
      b.tbin_hoff = h->tbin_hoff;
      b.bin_start = h->bin_start;
      b.bin_list = h->bin_list;
      b.wscale = h->bin_scale;
      void (*kb)(Bin3D<R>) = nullptr;
      if constexpr (sizeof(R) == 8) kb = k3d_bin_f64; else kb = k3d_bin_f32;
      hipLaunchKernelGGL(kb, dim3(h->T), dim3(256), h->bin_lds, h->stream, b);
      HIPCHECK(hipGetLastError());
    }
    if (kind == K_STEP && h->wave_cpw &&
        (h->step_kernel == 2 || (h->step_kernel == 0 && h->C >= h->wave_min_cells))) {
      auto kw = reinterpret_cast<void (*)(Step3D<R>)>(const_cast<void*>(wstep_kernel(h->rs, h->nv, h->nf, probe)));
      hipLaunchKernelGGL(kw, dim3((h->C + h->wave_cpw - 1) / h->wave_cpw), dim3(64 * h->wave_cpw), h->lds_w,
                         h->stream, a);
    } else {
      hipLaunchKernelGGL(k, dim3(h->C), dim3(h->NT), h->lds, h->stream, a);
    }
    HIPCHECK(hipGetLastError());
    return DPM_OK;
  };
  return h->rs == 8 ? go(0.0) : go(0.0f);
}

static int ensure_stats(dpm_tissue* h) {
  if (h->stats_valid) return DPM_OK;
  const int c = h->cur;
  TRY(launch3d(h, K_STATS, c, 0.0f, 0, [&](auto& a) {
    using S = std::remove_reference_t<decltype(*a.st_out)>;
    using R = std::remove_reference_t<decltype(*a.em_out)>;
    a.st_out = (S*)h->st[c];
    a.vol_out = h->vol[c];
    a.em_out = (R*)h->em[c];
  }));
  h->stats_valid = true;
  return DPM_OK;
}

static int launch_step3d(dpm_tissue* h, int in, float dt, int flags) {
  return launch3d(h, K_STEP, in, dt, flags, [](auto&) {});
}

// A graph of kGraphSteps steps starting from buffer 0 (ends on buffer 0).
static int build_graph3d(dpm_tissue* h, float dt) {
  if (h->graph) {
    (void)hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  hipGraph_t g = nullptr;
  HIPCHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int rc = DPM_OK;
  for (int s = 0; s < kGraphSteps && !rc; s++) rc = launch_step3d(h, s & 1, dt, F_MOVE);
  hipError_t e = hipStreamEndCapture(h->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(DPM_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&h->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    h->graph = nullptr;
    return fail(DPM_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  h->graph_dt = dt;
  return DPM_OK;
}

static int euler3d(dpm_tissue* h, int nsteps, float dt) {
  if (nsteps <= 0) return fail(DPM_EINVAL, "nsteps must be positive");
  if (!(dt > 0.0f) || dt > 0.1f) return fail(DPM_EINVAL, "dt must be positive and reasonable");
  TRY(ensure_stats(h));
  // the step graph for this dt is captured on the first call with it (a
  // capture launches nothing), so that later calls replay it
  if (!h->graph || h->graph_dt != dt) TRY(build_graph3d(h, dt));
  int left = nsteps - 1;  // the last step also writes the forces
  if (left >= kGraphSteps) {
    // Bring the ping-pong to buffer 0 so that the graph (captured from 0)
    // applies.
    if (h->cur != 0) {
      TRY(launch_step3d(h, h->cur, dt, F_MOVE));
      h->cur = 1 - h->cur;
      left--;
    }
    if (left >= kGraphSteps) {
      while (left >= kGraphSteps) {
        HIPCHECK(hipGraphLaunch(h->graph, h->stream));
        left -= kGraphSteps;
      }
    }
  }
  for (; left > 0; left--) {
    TRY(launch_step3d(h, h->cur, dt, F_MOVE));
    h->cur = 1 - h->cur;
  }
  TRY(launch_step3d(h, h->cur, dt, F_MOVE | F_WRITE_FORCES));
  h->cur = 1 - h->cur;
  h->forces_valid = true;
  return DPM_OK;
}

// Cells per workgroup of the wave form (1, 2 or 4, within 64 KiB of LDS).
static int set_wave_cpw(dpm_tissue* h, int cpw) {
  const size_t lds = ldsw_layout(h->nv, h->nf, cpw, h->rs).total;
  if (lds > 65536) return fail(DPM_ERANGE, "cells per workgroup exceed the LDS of a workgroup");
  h->wave_cpw = cpw;
  h->lds_w = lds;
  for (bool probe : {false, true})
    HIPCHECK(hipFuncSetAttribute(wstep_kernel(h->rs, h->nv, h->nf, probe), hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)lds));
  if (h->graph) {
    (void)hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  return DPM_OK;
}

extern "C" int dpm3d_autotune(dpm_tissue* h, int steps, float dt, int* best, float* times_ms) {
  if (!h || h->dim != 3) return fail(DPM_EINVAL, "not a 3D tissue");
  if (steps < 1 || steps > 1000) return fail(DPM_EINVAL, "steps must be in 1..1000");
  if (!(dt > 0.0f) || dt > 0.1f) return fail(DPM_EINVAL, "dt must be positive and reasonable");
  TRY(ensure_stats(h));
  // the state the trials start from, restored after each of them
  const int c0 = h->cur;
  const size_t NVT = (size_t)h->C * h->nv, rs = (size_t)h->rs, C = (size_t)h->C;
  const size_t sz_pos = NVT * rs, sz_lo = h->lo ? 3 * NVT * rs : 0, sz_st = C * kStatWords * rs,
               sz_vol = C * sizeof(double), sz_em = C * rs, sz_cnt = 4 * DPM_COUNTER_SLOTS * sizeof(unsigned long long);
  const size_t total = 3 * sz_pos + sz_lo + sz_st + sz_vol + sz_em + sz_cnt;
  unsigned char* save = nullptr;
  TRY(dev_alloc(&save, total));
  struct Piece { void* dev; size_t n; };
  const Piece pieces[] = {{h->x[c0], sz_pos}, {h->y[c0], sz_pos}, {h->z[c0], sz_pos}, {h->lo, sz_lo},
                          {h->st[c0], sz_st}, {h->vol[c0], sz_vol}, {h->em[c0], sz_em}, {h->counters, sz_cnt}};
  auto copy_all = [&](bool to_save) -> int {
    size_t off = 0;
    for (const Piece& q : pieces) {
      if (q.n && hipMemcpyAsync(to_save ? save + off : q.dev, to_save ? q.dev : save + off, q.n,
                                hipMemcpyDeviceToDevice, h->stream) != hipSuccess)
        return fail(DPM_EDEVICE, "state copy failed");
      off += q.n;
    }
    h->cur = c0;
    return DPM_OK;
  };
  int rc = copy_all(true);
  // configurations: {form, cells per workgroup, cell list}, in the order of times_ms
  const int cfg[8][3] = {{1, 0, 0}, {1, 0, 1}, {2, 1, 0}, {2, 2, 0}, {2, 4, 0}, {2, 1, 1}, {2, 2, 1}, {2, 4, 1}};
  const int kind0 = h->step_kernel, cpw0 = h->wave_cpw, mode0 = h->cell_list_mode;
  const bool cells0 = h->cells_on;
  // the cell-list configurations each take the fastest of three bin widths
  const float scales[3] = {1.0f, 1.5f, 2.0f};
  const float scale0 = h->bin_scale;
  float sbest = scale0;
  float tbest = 0.0f;
  int ib = -1;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(DPM_EDEVICE, "hipEventCreate failed");
  for (int i = 0; i < 8 && !rc; i++) {
    if (times_ms) times_ms[i] = -1.0f;
    if (cfg[i][0] == 2 && (!h->wave_cpw || ldsw_layout(h->nv, h->nf, cfg[i][1], h->rs).total > 65536)) continue;
    h->step_kernel = cfg[i][0];
    if (cfg[i][0] == 2) rc = set_wave_cpw(h, cfg[i][1]);
    h->cells_on = cfg[i][2] == 1;
    const int nscale = cfg[i][2] == 1 ? 3 : 1;
    for (int si = 0; si < nscale && !rc; si++) {
      h->bin_scale = cfg[i][2] == 1 ? scales[si] : scale0;
      for (int rep = 0; rep < 2 && !rc; rep++) {  // a warm-up launch, then the timed steps
        rc = copy_all(false);
        if (rc) break;
        if (rep == 0) {
          rc = launch_step3d(h, h->cur, dt, F_MOVE);
          continue;
        }
        if (hipEventRecord(e0, h->stream) != hipSuccess) rc = fail(DPM_EDEVICE, "hipEventRecord failed");
        for (int k = 0; k < steps && !rc; k++) {
          rc = launch_step3d(h, h->cur, dt, F_MOVE);
          h->cur = 1 - h->cur;
        }
        if (!rc && (hipEventRecord(e1, h->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
          rc = fail(DPM_EDEVICE, "timing failed");
        float ms = 0.0f;
        if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = fail(DPM_EDEVICE, "timing failed");
        if (!rc) {
          const float per = ms / (float)steps;
          if (times_ms && (times_ms[i] < 0.0f || per < times_ms[i])) times_ms[i] = per;
          if (ib < 0 || ms < tbest) {
            tbest = ms;
            ib = i;
            sbest = h->bin_scale;
          }
        }
      }
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  int rc2 = copy_all(false);
  if (!rc && rc2) rc = rc2;
  if (hipStreamSynchronize(h->stream) != hipSuccess && !rc) rc = fail(DPM_EDEVICE, "stream sync failed");
  (void)hipFree(save);
  if (rc || ib < 0) {  // leave the handle as it was
    h->step_kernel = kind0;
    if (cpw0) (void)set_wave_cpw(h, cpw0);
    h->cells_on = cells0;
    h->cell_list_mode = mode0;
    h->bin_scale = scale0;
    return rc ? rc : fail(DPM_EDEVICE, "no configuration ran");
  }
  h->bin_scale = sbest;
  h->step_kernel = cfg[ib][0];
  if (cfg[ib][0] == 2) TRY(set_wave_cpw(h, cfg[ib][1]));
  h->cells_on = cfg[ib][2] == 1;
  h->cell_list_mode = cfg[ib][2];
  if (h->graph) {
    (void)hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  if (best) {
    best[0] = cfg[ib][0];
    best[1] = cfg[ib][0] == 2 ? cfg[ib][1] : 0;
    best[2] = cfg[ib][2];
  }
  return DPM_OK;
}

extern "C" int dpm3d_cell_stats(dpm_tissue* h, double* volume, float* com) {
  if (!h || h->dim != 3) return fail(DPM_EINVAL, "not a 3D tissue");
  TRY(ensure_stats(h));
  const int C = h->C;
  std::vector<unsigned char> st((size_t)C * kStatWords * h->rs);
  HIPCHECK(hipMemcpyAsync(st.data(), h->st[h->cur], st.size(), hipMemcpyDeviceToHost, h->stream));
  if (volume)
    HIPCHECK(hipMemcpyAsync(volume, h->vol[h->cur], C * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  if (com)
    for (int c = 0; c < C; c++)
      for (int k = 0; k < 3; k++)
        com[3 * c + k] = h->rs == 8 ? (float)((const double*)st.data())[kStatWords * c + k]
                                    : ((const float*)st.data())[kStatWords * c + k];
  return DPM_OK;
}

extern "C" int dpm3d_cell_measures(dpm_tissue* h, float* volume, float* surface_area, float* com) {
  if (!h || h->dim != 3) return fail(DPM_EINVAL, "not a 3D tissue");
  const size_t lds = measures_lds_bytes(h->nv, h->nf);
  if ((int)lds > max_lds_per_block(h->device)) return fail(DPM_ERANGE, "cell mesh does not fit the LDS of one workgroup");
  if (!h->meas) TRY(h->alloc(&h->meas, 5 * (size_t)h->C));
  const int c = h->cur;
  auto go = [&](auto tag) -> int {
    using R = decltype(tag);
    Measures3D<R> a{(const R*)h->x[c], (const R*)h->y[c], (const R*)h->z[c], h->faces, h->nv, h->nf, h->meas};
    void (*k)(Measures3D<R>) = nullptr;
    if constexpr (sizeof(R) == 8) k = k3d_measures_f64; else k = k3d_measures_f32;
    HIPCHECK(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k, dim3(h->C), dim3(64), lds, h->stream, a);
    HIPCHECK(hipGetLastError());
    return DPM_OK;
  };
  TRY(h->rs == 8 ? go(0.0) : go(0.0f));
  std::vector<float> m(5 * (size_t)h->C);
  HIPCHECK(hipMemcpyAsync(m.data(), h->meas, m.size() * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  for (int i = 0; i < h->C; i++) {
    if (volume) volume[i] = m[5 * (size_t)i];
    if (surface_area) surface_area[i] = m[5 * (size_t)i + 1];
    if (com)
      for (int k = 0; k < 3; k++) com[3 * (size_t)i + k] = m[5 * (size_t)i + 2 + k];
  }
  return DPM_OK;
}

extern "C" int dpm3d_energy(dpm_tissue* h, double* energy) {
  if (!h || h->dim != 3 || !energy) return fail(DPM_EINVAL, "bad arguments");
  TRY(launch3d(h, K_ENERGY, h->cur, 0.0f, 0, [](auto&) {}));
  HIPCHECK(hipMemcpyAsync(energy, h->energy, h->C * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  return DPM_OK;
}

extern "C" int dpm3d_probe(dpm_tissue* h, float* forces, unsigned char* contact, float* winding,
                           unsigned char* down, int* adhesion) {
  if (!h || h->dim != 3) return fail(DPM_EINVAL, "not a 3D tissue");
  TRY(ensure_stats(h));
  const size_t NVT = (size_t)h->C * h->nv;
  const size_t npair = NVT * (size_t)h->ncmax;
  unsigned char *dc = nullptr, *dd = nullptr;
  float* dw = nullptr;
  int* da = nullptr;
  hipStream_t s = h->stream;
  int rc = DPM_OK;
  if (contact) rc |= dev_alloc(&dc, npair);
  if (winding) rc |= dev_alloc(&dw, npair);
  if (down) rc |= dev_alloc(&dd, (size_t)h->C * h->nf);
  if (adhesion) rc |= dev_alloc(&da, NVT);
  if (!rc) {
    if (dc) (void)hipMemsetAsync(dc, 0, npair, s);
    if (dw) (void)hipMemsetAsync(dw, 0, npair * sizeof(float), s);
    if (dd) (void)hipMemsetAsync(dd, 0, (size_t)h->C * h->nf, s);
    if (da) (void)hipMemsetAsync(da, 0, NVT * sizeof(int), s);
    rc = launch3d(h, K_STEP, h->cur, 0.0f, F_WRITE_FORCES, [&](auto& a) {
      a.counters = nullptr;
      a.p_contact = dc;
      a.p_winding = dw;
      a.p_down = dd;
      a.p_adh = da;
    });
  }
  if (!rc && dc) rc = hipMemcpyAsync(contact, dc, npair, hipMemcpyDeviceToHost, s) ? DPM_EDEVICE : 0;
  if (!rc && dw)
    rc = hipMemcpyAsync(winding, dw, npair * sizeof(float), hipMemcpyDeviceToHost, s) ? DPM_EDEVICE : 0;
  if (!rc && dd)
    rc = hipMemcpyAsync(down, dd, (size_t)h->C * h->nf, hipMemcpyDeviceToHost, s) ? DPM_EDEVICE : 0;
  if (!rc && da)
    rc = hipMemcpyAsync(adhesion, da, NVT * sizeof(int), hipMemcpyDeviceToHost, s) ? DPM_EDEVICE : 0;
  if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = fail(DPM_EDEVICE, "stream sync failed");
  if (dc) (void)hipFree(dc);
  if (dw) (void)hipFree(dw);
  if (dd) (void)hipFree(dd);
  if (da) (void)hipFree(da);
  if (rc) return rc;
  h->forces_valid = true;
  if (forces) TRY(dpm3d_get_forces(h, forces, NVT));
  return DPM_OK;
}

extern "C" int dpm3d_probe_adhesion(dpm_tissue* h, unsigned* own_keys, unsigned* rev_keys) {
  if (!h || h->dim != 3) return fail(DPM_EINVAL, "not a 3D tissue");
  if (!own_keys && !rev_keys) return DPM_OK;
  TRY(ensure_stats(h));
  const size_t NVT = (size_t)h->C * h->nv;
  unsigned* dk = nullptr;
  hipStream_t s = h->stream;
  int rc = dev_alloc(&dk, 2 * NVT);
  if (!rc) {
    (void)hipMemsetAsync(dk, 0, 2 * NVT * sizeof(unsigned), s);
    rc = launch3d(h, K_STEP, h->cur, 0.0f, F_WRITE_FORCES, [&](auto& a) {
      a.counters = nullptr;
      a.p_adh_hash = dk;
    });
  }
  std::vector<unsigned> k(rc ? 0 : 2 * NVT);
  if (!rc)
    rc = hipMemcpyAsync(k.data(), dk, 2 * NVT * sizeof(unsigned), hipMemcpyDeviceToHost, s) ? DPM_EDEVICE : 0;
  if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = fail(DPM_EDEVICE, "stream sync failed");
  if (dk) (void)hipFree(dk);
  if (rc) return rc;
  h->forces_valid = true;
  for (size_t i = 0; i < NVT; i++) {
    if (own_keys) own_keys[i] = k[2 * i];
    if (rev_keys) rev_keys[i] = k[2 * i + 1];
  }
  return DPM_OK;
}

extern "C" int dpm3d_probe_contact(dpm_tissue* h, unsigned* count, unsigned* keys) {
  if (!h || h->dim != 3) return fail(DPM_EINVAL, "not a 3D tissue");
  if (!count && !keys) return DPM_OK;
  TRY(ensure_stats(h));
  const size_t NVT = (size_t)h->C * h->nv;
  unsigned* dk = nullptr;
  hipStream_t s = h->stream;
  int rc = dev_alloc(&dk, 2 * NVT);
  if (!rc) {
    (void)hipMemsetAsync(dk, 0, 2 * NVT * sizeof(unsigned), s);
    rc = launch3d(h, K_STEP, h->cur, 0.0f, F_WRITE_FORCES, [&](auto& a) {
      a.counters = nullptr;
      a.p_contact_key = dk;
    });
  }
  std::vector<unsigned> k(rc ? 0 : 2 * NVT);
  if (!rc)
    rc = hipMemcpyAsync(k.data(), dk, 2 * NVT * sizeof(unsigned), hipMemcpyDeviceToHost, s) ? DPM_EDEVICE : 0;
  if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = fail(DPM_EDEVICE, "stream sync failed");
  if (dk) (void)hipFree(dk);
  if (rc) return rc;
  h->forces_valid = true;
  for (size_t i = 0; i < NVT; i++) {
    if (count) count[i] = k[2 * i];
    if (keys) keys[i] = k[2 * i + 1];
  }
  return DPM_OK;
}

// --------------------------------------------- probe without allocation
// dpm3d_probe_noalloc keeps its device tables (pair flags, winding numbers,
// facing-down flags, adhesion counts, the AoS force copy) in this array
// instead of allocating them per call.  Each device holds its own copy at
// its own address, resolved once per device by hipGetSymbolAddress; host
// code uses only that address, never the array's name.  Every call on a
// device shares the device's copy (calls that overlap on one device race).
constexpr size_t kProbeScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kProbeScratchBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int scratch_address(int dev, unsigned char** out) {
  *out = nullptr;
  if (dev < 0 || dev >= kMaxDevices) return fail(DPM_ERANGE, "device index beyond the probe scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int prev = 0;
    HIPCHECK(hipGetDevice(&prev));
    HIPCHECK(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    (void)hipSetDevice(prev);
    if (e != hipSuccess || !p)
      return fail(DPM_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = static_cast<unsigned char*>(g_scratch_ptr[dev]);
  return DPM_OK;
}
}  // namespace

__global__ void k_fill_zero(unsigned* p, size_t nwords) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nwords; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

template <typename R>
__global__ void k_soa_to_aos_f32(const R* x, const R* y, const R* z, float* out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    out[3 * i] = (float)x[i];
    out[3 * i + 1] = (float)y[i];
    out[3 * i + 2] = (float)z[i];
  }
}

extern "C" int dpm3d_probe_noalloc(dpm_tissue* h, float* forces, unsigned char* contact, float* winding,
                                   unsigned char* down, int* adhesion) {
  if (!h || h->dim != 3) return fail(DPM_EINVAL, "not a 3D tissue");
  const size_t NVT = (size_t)h->C * h->nv;
  const size_t npair = NVT * (size_t)h->ncmax;
  auto al = [](size_t b) { return (b + 255u) & ~(size_t)255u; };
  // table offsets in the scratch, each 256-byte aligned and a multiple of 4 bytes
  const size_t o_contact = 0;
  const size_t o_winding = o_contact + (contact ? al(npair) : 0);
  const size_t o_down = o_winding + (winding ? al(npair * sizeof(float)) : 0);
  const size_t o_adh = o_down + (down ? al((size_t)h->C * h->nf) : 0);
  const size_t o_forces = o_adh + (adhesion ? al(NVT * sizeof(int)) : 0);
  const size_t total = o_forces + (forces ? al(3 * NVT * sizeof(float)) : 0);
  if (total > kProbeScratchBytes) {
    char buf[200];
    snprintf(buf, sizeof(buf),
             "dpm3d_probe_noalloc: the requested tables need %zu bytes, more than its %zu-byte device scratch "
             "(use dpm3d_probe, or request fewer tables)",
             total, kProbeScratchBytes);
    return fail(DPM_ERANGE, buf);
  }
  unsigned char* base = nullptr;
  TRY(scratch_address(h->device, &base));
  hipStream_t s = h->stream;
  unsigned char* dc = contact ? base + o_contact : nullptr;
  float* dw = winding ? reinterpret_cast<float*>(base + o_winding) : nullptr;
  unsigned char* dd = down ? base + o_down : nullptr;
  int* da = adhesion ? reinterpret_cast<int*>(base + o_adh) : nullptr;
  float* df = forces ? reinterpret_cast<float*>(base + o_forces) : nullptr;
  if (o_forces > 0) {  // the tables start at zero (the array keeps the last call's contents)
    const size_t words = o_forces / 4;
    k_fill_zero<<<(int)std::min<size_t>((words + 255) / 256, 4096), 256, 0, s>>>(reinterpret_cast<unsigned*>(base),
                                                                                 words);
    HIPCHECK(hipGetLastError());
  }
  TRY(ensure_stats(h));
  TRY(launch3d(h, K_STEP, h->cur, 0.0f, F_WRITE_FORCES, [&](auto& a) {
    a.counters = nullptr;
    a.p_contact = dc;
    a.p_winding = dw;
    a.p_down = dd;
    a.p_adh = da;
  }));
  h->forces_valid = true;
  if (df) {
    const int blocks = (int)std::min<size_t>((NVT + 255) / 256, 4096);
    if (h->rs == 8)
      k_soa_to_aos_f32<double><<<blocks, 256, 0, s>>>((const double*)h->fx, (const double*)h->fy,
                                                      (const double*)h->fz, df, NVT);
    else
      k_soa_to_aos_f32<float><<<blocks, 256, 0, s>>>((const float*)h->fx, (const float*)h->fy,
                                                     (const float*)h->fz, df, NVT);
    HIPCHECK(hipGetLastError());
  }
  if (dc) HIPCHECK(hipMemcpyAsync(contact, dc, npair, hipMemcpyDeviceToHost, s));
  if (dw) HIPCHECK(hipMemcpyAsync(winding, dw, npair * sizeof(float), hipMemcpyDeviceToHost, s));
  if (dd) HIPCHECK(hipMemcpyAsync(down, dd, (size_t)h->C * h->nf, hipMemcpyDeviceToHost, s));
  if (da) HIPCHECK(hipMemcpyAsync(adhesion, da, NVT * sizeof(int), hipMemcpyDeviceToHost, s));
  if (df) HIPCHECK(hipMemcpyAsync(forces, df, 3 * NVT * sizeof(float), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  return DPM_OK;
}

// ------------------------------------------------------------- 3D adjoint

template <typename In>
static int upload_real3(dpm_tissue* h, void* dst, const In* src, size_t n) {
  if (h->rs == 8) {
    std::vector<double> v(src, src + n);
    TRY(upload((double*)dst, v.data(), n, h->stream));
  } else {
    std::vector<float> v(n);
    for (size_t i = 0; i < n; i++) v[i] = (float)src[i];
    TRY(upload((float*)dst, v.data(), n, h->stream));
  }
  HIPCHECK(hipStreamSynchronize(h->stream));
  return DPM_OK;
}

static int ensure_real(dpm_tissue* h, void** p, size_t& cap, size_t n) {
  if (*p && cap >= n) return DPM_OK;
  if (*p) {
    (void)hipFree(*p);
    h->allocs.erase(std::remove(h->allocs.begin(), h->allocs.end(), *p), h->allocs.end());
    *p = nullptr;
  }
  TRY(h->alloc_real(p, n));
  cap = n;
  return DPM_OK;
}

static int copy_positions_dd(dpm_tissue* h, void* dx, void* dy, void* dz, const void* sx, const void* sy,
                             const void* sz) {
  const size_t b = (size_t)h->C * h->nv * h->rs;
  HIPCHECK(hipMemcpyAsync(dx, sx, b, hipMemcpyDeviceToDevice, h->stream));
  HIPCHECK(hipMemcpyAsync(dy, sy, b, hipMemcpyDeviceToDevice, h->stream));
  HIPCHECK(hipMemcpyAsync(dz, sz, b, hipMemcpyDeviceToDevice, h->stream));
  return DPM_OK;
}

static void* slot(const dpm_tissue* h, void* base, size_t k) {
  return (unsigned char*)base + k * (size_t)h->C * h->nv * h->rs;
}

extern "C" int dpm3d_trajectory(dpm_tissue* h, int nsteps, float dt, int every) {
  if (!h || h->dim != 3) return fail(DPM_EINVAL, "not a 3D tissue");
  if (nsteps <= 0 || every <= 0 || nsteps % every) return fail(DPM_EINVAL, "nsteps must be a positive multiple of every");
  if (!(dt > 0.0f) || dt > 0.1f) return fail(DPM_EINVAL, "dt must be positive and reasonable");
  if (h->mask & DPM3D_ADHESION)
    return fail(DPM_EINVAL, "the adjoint does not include vertex adhesion: clear DPM3D_ADHESION");
  if ((int)h->lds_vjp > max_lds_per_block(h->device))
    return fail(DPM_ERANGE, "cell mesh does not fit the LDS of the adjoint kernel");
  const size_t NVT = (size_t)h->C * h->nv;
  const int nck = nsteps / every + 1;
  const int ns = h->lo ? 6 : 3;  // x, y, z (+ their residuals when compensated)
  TRY(ensure_real(h, &h->ck, h->ck_cap, (size_t)nck * ns * NVT));
  TRY(ensure_stats(h));
  for (int k = 0; k < nck; k++) {
    void* c = slot(h, h->ck, (size_t)ns * k);
    TRY(copy_positions_dd(h, c, slot(h, c, 1), slot(h, c, 2), h->x[h->cur], h->y[h->cur], h->z[h->cur]));
    if (h->lo)
      HIPCHECK(hipMemcpyAsync(slot(h, c, 3), h->lo, 3 * NVT * h->rs, hipMemcpyDeviceToDevice, h->stream));
    if (k < nck - 1) TRY(euler3d(h, every, dt));
  }
  HIPCHECK(hipStreamSynchronize(h->stream));
  h->traj_slots = ns;
  h->traj_steps = nsteps;
  h->traj_every = every;
  h->traj_dt = dt;
  return DPM_OK;
}

template <typename R>
static Vjp3D<R> vjp_args(dpm_tissue* h, size_t s) {
  Vjp3D<R> a{};
  void* xs = slot(h, h->seg, 3 * s);
  a.x = (const R*)xs;
  a.y = (const R*)slot(h, xs, 1);
  a.z = (const R*)slot(h, xs, 2);
  a.w = (const R*)slot(h, h->segw, s);
  const size_t NP = (size_t)h->C * h->nv * DPM3D_MAXPART;
  a.rec_cell = h->segcell + s * NP;
  a.rec_w = (const R*)h->segwp + s * NP;
  a.rec_shift = (const R*)h->segsh + s * NP * 3;
  a.cx = (R*)h->cross;
  a.cy = (R*)slot(h, h->cross, 1);
  a.cz = (R*)slot(h, h->cross, 2);
  a.lx = (R*)h->lam;
  a.ly = (R*)slot(h, h->lam, 1);
  a.lz = (R*)slot(h, h->lam, 2);
  a.gtheta = h->gtheta;
  a.Kv = h->Kv;
  a.Ka = h->Ka;
  a.Ks = h->Ks;
  a.Kb = h->Kb;
  a.v0 = h->v0;
  a.a0 = h->a0;
  a.cell_tissue = h->cell_tissue;
  a.tL = h->tL;
  a.tKre = h->tKre;
  a.tksg = h->tksg;
  a.faces = h->faces;
  a.vf = h->vf;
  a.adj = h->adj;
  a.nv = h->nv;
  a.nf = h->nf;
  a.dt = (R)h->traj_dt;
  a.mask = h->mask;
  return a;
}

// lambda_final and lambda0 are AoS (reference order) in In/Out precision.
template <typename In, typename Out>
static int adjoint3d(dpm_tissue* h, const In* lambda_final, double* dtheta, Out* lambda0) {
  if (!h || h->dim != 3 || !lambda_final) return fail(DPM_EINVAL, "bad arguments");
  if (!h->traj_steps) return fail(DPM_EINVAL, "no recorded trajectory: call dpm3d_trajectory first");
  if ((h->lo ? 6 : 3) != h->traj_slots)
    return fail(DPM_EINVAL, "compensation changed since the trajectory was recorded");
  const int ns = h->traj_slots;
  if (h->mask & DPM3D_ADHESION)
    return fail(DPM_EINVAL, "the adjoint does not include vertex adhesion: clear DPM3D_ADHESION");
  const size_t NVT = (size_t)h->C * h->nv;
  const int every = h->traj_every, nseg = h->traj_steps / every;
  TRY(ensure_real(h, &h->seg, h->seg_cap, (size_t)every * 3 * NVT));
  TRY(ensure_real(h, &h->segw, h->segw_cap, (size_t)every * NVT));
  TRY(ensure_real(h, &h->lam, h->lam_cap, 3 * NVT));
  TRY(ensure_real(h, &h->segwp, h->segwp_cap, (size_t)every * NVT * DPM3D_MAXPART));
  TRY(ensure_real(h, &h->segsh, h->segsh_cap, (size_t)every * NVT * DPM3D_MAXPART * 3));
  TRY(ensure_real(h, &h->cross, h->cross_cap, 3 * NVT));
  if (h->segcell_cap < (size_t)every * NVT * DPM3D_MAXPART) {
    if (h->segcell) {
      (void)hipFree(h->segcell);
      h->allocs.erase(std::remove(h->allocs.begin(), h->allocs.end(), (void*)h->segcell), h->allocs.end());
      h->segcell = nullptr;
    }
    TRY(h->alloc(&h->segcell, (size_t)every * NVT * DPM3D_MAXPART));
    h->segcell_cap = (size_t)every * NVT * DPM3D_MAXPART;
  }
  if (!h->gtheta) TRY(h->alloc(&h->gtheta, (size_t)h->C * 4));
  if (!h->rec_overflow) TRY(h->alloc(&h->rec_overflow, 1));
  HIPCHECK(hipMemsetAsync(h->rec_overflow, 0, sizeof(int), h->stream));
  HIPCHECK(hipMemsetAsync(h->cross, 0, 3 * NVT * h->rs, h->stream));
  // lambda_N
  {
    std::vector<In> a(NVT), b(NVT), c(NVT);
    for (size_t i = 0; i < NVT; i++) {
      a[i] = lambda_final[3 * i];
      b[i] = lambda_final[3 * i + 1];
      c[i] = lambda_final[3 * i + 2];
    }
    TRY(upload_real3(h, h->lam, a.data(), NVT));
    TRY(upload_real3(h, slot(h, h->lam, 1), b.data(), NVT));
    TRY(upload_real3(h, slot(h, h->lam, 2), c.data(), NVT));
  }
  HIPCHECK(hipMemsetAsync(h->gtheta, 0, sizeof(double) * h->C * 4, h->stream));
  const float dt = h->traj_dt;
  for (int sg = nseg - 1; sg >= 0; sg--) {
    // recompute the segment from its checkpoint, keeping x_n and W_n
    void* c = slot(h, h->ck, (size_t)ns * sg);
    h->cur = 0;
    TRY(copy_positions_dd(h, h->x[0], h->y[0], h->z[0], c, slot(h, c, 1), slot(h, c, 2)));
    if (h->lo)
      HIPCHECK(hipMemcpyAsync(h->lo, slot(h, c, 3), 3 * NVT * h->rs, hipMemcpyDeviceToDevice, h->stream));
    h->stats_valid = false;
    TRY(ensure_stats(h));
    for (int s = 0; s < every; s++) {
      void* xs = slot(h, h->seg, 3 * (size_t)s);
      TRY(copy_positions_dd(h, xs, slot(h, xs, 1), slot(h, xs, 2), h->x[h->cur], h->y[h->cur], h->z[h->cur]));
      void* w = slot(h, h->segw, s);
      const size_t NP = NVT * DPM3D_MAXPART;
      TRY(launch3d(h, K_STEP, h->cur, dt, F_MOVE, [&](auto& a) {
        using R = std::remove_reference_t<decltype(*a.w_out)>;
        a.w_out = (R*)w;
        a.rec_cell = h->segcell + s * NP;
        a.rec_w = (R*)h->segwp + s * NP;
        a.rec_shift = (R*)h->segsh + s * NP * 3;
        a.rec_overflow = h->rec_overflow;
      }));
      h->cur = 1 - h->cur;
    }
    // reverse sweep through the segment
    for (int s = every - 1; s >= 0; s--) {
      const int cb = (int)std::min<size_t>((NVT + 255) / 256, 1024);
      if (h->rs == 8) {
        Vjp3D<double> a = vjp_args<double>(h, s);
        if (h->NT == 256)
          hipLaunchKernelGGL(k3d_vjp_f64_t256, dim3(h->C), dim3(h->NT), h->lds_vjp, h->stream, a);
        else
          hipLaunchKernelGGL(k3d_vjp_f64, dim3(h->C), dim3(h->NT), h->lds_vjp, h->stream, a);
        hipLaunchKernelGGL(k3d_vjp_cross_f64, dim3(cb), dim3(256), 0, h->stream, a.lx, a.ly, a.lz, a.cx, a.cy,
                           a.cz, (double)a.dt, NVT);
      } else {
        Vjp3D<float> a = vjp_args<float>(h, s);
        hipLaunchKernelGGL(k3d_vjp_f32, dim3(h->C), dim3(h->NT), h->lds_vjp, h->stream, a);
        hipLaunchKernelGGL(k3d_vjp_cross_f32, dim3(cb), dim3(256), 0, h->stream, a.lx, a.ly, a.lz, a.cx, a.cy,
                           a.cz, a.dt, NVT);
      }
      HIPCHECK(hipGetLastError());
    }
  }
  // back to the final state of the trajectory
  void* last = slot(h, h->ck, (size_t)ns * nseg);
  h->cur = 0;
  TRY(copy_positions_dd(h, h->x[0], h->y[0], h->z[0], last, slot(h, last, 1), slot(h, last, 2)));
  if (h->lo)
    HIPCHECK(hipMemcpyAsync(h->lo, slot(h, last, 3), 3 * NVT * h->rs, hipMemcpyDeviceToDevice, h->stream));
  h->stats_valid = false;
  if (dtheta)
    HIPCHECK(hipMemcpyAsync(dtheta, h->gtheta, sizeof(double) * h->C * 4, hipMemcpyDeviceToHost, h->stream));
  if (lambda0) TRY(download3(h, h->lam, slot(h, h->lam, 1), slot(h, h->lam, 2), lambda0, NVT));
  int overflow = 0;
  HIPCHECK(hipMemcpyAsync(&overflow, h->rec_overflow, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  if (overflow)
    return fail(DPM_ERANGE, "a vertex was in contact with more than 4 cells: the adjoint does not cover it");
  return DPM_OK;
}

extern "C" int dpm3d_adjoint(dpm_tissue* h, const float* lambda_final, double* dtheta, float* lambda0) {
  return adjoint3d(h, lambda_final, dtheta, lambda0);
}
extern "C" int dpm3d_adjoint_f64(dpm_tissue* h, const double* lambda_final, double* dtheta, double* lambda0) {
  return adjoint3d(h, lambda_final, dtheta, lambda0);
}

// ================================================================== 2D ====

static_assert(sizeof(dpm_relax_stats) == RS_COUNT * sizeof(double), "dpm_relax_stats layout");

static int create2d(const dpm2d_desc* d, dpm_tissue** out, int rs) {
  if (!d || !out) return fail(DPM_EINVAL, "null argument");
  *out = nullptr;
  const int T = d->ntissues;
  if (T <= 0 || !d->cell_begin || d->cell_begin[0] != 0) return fail(DPM_EINVAL, "bad cell_begin");
  const int C = d->cell_begin[T];
  if (!d->nv || !d->Ka || !d->Kl || !d->Kb || !d->a0 || !d->l0 || !d->r0 || !d->L || !d->pbc ||
      !d->Kre || !d->Kat)
    return fail(DPM_EINVAL, "missing parameter array");
  std::vector<int> cvb(C + 1, 0);
  for (int c = 0; c < C; c++) {
    if (d->nv[c] < 3) return fail(DPM_EINVAL, "a cell needs at least 3 vertices");
    cvb[c + 1] = cvb[c] + d->nv[c];
  }
  int maxv = 0, maxc = 0;
  for (int t = 0; t < T; t++) {
    const int nc = d->cell_begin[t + 1] - d->cell_begin[t];
    if (nc <= 0) return fail(DPM_EINVAL, "NCELLS must be positive");
    if (nc > 32767) return fail(DPM_ERANGE, "more than 32767 cells in one tissue");
    maxc = std::max(maxc, nc);
    maxv = std::max(maxv, cvb[d->cell_begin[t + 1]] - cvb[d->cell_begin[t]]);
  }
  dpm_tissue* h = new dpm_tissue();
  h->dim = 2;
  h->rs = rs;
  int rc = init_common(h);
  if (rc) { delete h; return rc; }
  h->T = T;
  h->C = C;
  h->V = cvb[C];
  // Four threads per vertex (up to 1024) when some tissue can take the
  // crossing-parity pass of k2d_run (its split_over predicate at that size;
  // the mask is left out, as dpm2d_set_tissue_params may replace it), which
  // spreads (vertex, candidate) pairs over all of them; otherwise one per
  // vertex and DPM2D_CSPLIT per cell for the candidate pass.
  const int nt4 = std::min(1024, std::max(64, (4 * maxv + 63) / 64 * 64));
  bool any_pass = false;
  for (int t = 0; t < T; t++) {
    const int nc = d->cell_begin[t + 1] - d->cell_begin[t];
    const int nvt = cvb[d->cell_begin[t + 1]] - cvb[d->cell_begin[t]];
    if (nvt <= 32 * DPM2D_CSPLIT * nc && (long long)nvt * nc <= 16LL * nt4)
      any_pass = true;
  }
  h->NT2 = any_pass ? nt4 : std::min(1024, std::max(64, (std::max(maxv, DPM2D_CSPLIT * maxc) + 63) / 64 * 64));
  h->NTr = std::min(1024, std::max(64, (maxv / 4 + 63) / 64 * 64));
  h->lds2 = lds2d_bytes(maxv, maxc, rs);
  // a tissue beyond one workgroup (LDS, or 16 vertices per thread) takes the
  // wide path; float tissues only
  h->fits_wg = maxv <= DPM2D_VPT * h->NT2 && (int)h->lds2 <= max_lds_per_block(h->device);
  if (!h->fits_wg && rs == 8) {
    delete h;
    return fail(DPM_ERANGE, "a double 2D tissue must fit the LDS of one workgroup");
  }
  if (!h->fits_wg) h->lds2 = 0;
  // the crossing-parity pair list where a tissue can take the pass and the
  // LDS allows it (without it the pass tests every pair where it finds it)
  h->lds_run = h->lds2;
  h->over_cap = 0;
  if (h->fits_wg && any_pass && (int)(h->lds2 + lds2d_over_bytes()) <= max_lds_per_block(h->device)) {
    h->lds_run = h->lds2 + lds2d_over_bytes();
    h->over_cap = DPM2D_OVER_CAP;
  }
  h->max_cells = maxc;
  h->cvb_h = cvb;
  h->tissue_begin_h.assign(d->cell_begin, d->cell_begin + T + 1);
  auto AR = [&](void** p, size_t n) { return h->alloc_real(p, n); };
  rc = 0;
  rc |= AR(&h->x2, h->V);
  rc |= AR(&h->y2, h->V);
  rc |= AR(&h->fx2, h->V);
  rc |= AR(&h->fy2, h->V);
  rc |= AR(&h->tLr, T);
  rc |= h->alloc(&h->res2, (size_t)T * RS_COUNT);
  rc |= h->alloc(&h->sum_a0, T);
  rc |= h->alloc(&h->tcb, T + 1);
  rc |= h->alloc(&h->cvb, C + 1);
  rc |= h->alloc(&h->Ka, C);
  rc |= h->alloc(&h->Kl, C);
  rc |= h->alloc(&h->Kb, C);
  rc |= h->alloc(&h->a0, C);
  rc |= h->alloc(&h->l0_2, C);
  rc |= h->alloc(&h->r0, C);
  rc |= h->alloc(&h->tpbc, T);
  rc |= h->alloc(&h->tKre, T);
  rc |= h->alloc(&h->tKat, T);
  if (rc) { delete h; return DPM_ENOMEM; }
  hipStream_t s = h->stream;
  rc |= upload(h->tcb, d->cell_begin, T + 1, s);
  rc |= upload(h->cvb, cvb.data(), C + 1, s);
  (void)hipMemsetAsync(h->fx2, 0, (size_t)h->V * rs, s);
  (void)hipMemsetAsync(h->fy2, 0, (size_t)h->V * rs, s);
  (void)hipMemsetAsync(h->res2, 0, sizeof(double) * T * RS_COUNT, s);
  if (!rc) rc = dpm2d_set_cell_params(h, d->Ka, d->Kl, d->Kb, d->a0, d->l0, d->r0);
  if (!rc) rc = dpm2d_set_tissue_params(h, d->L, d->pbc, d->Kre, d->Kat, d->mask);
  if (rc) { delete h; return rc; }
  if (h->fits_wg) {
    (void)hipFuncSetAttribute((const void*)k2d_run, hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->lds_run);
    (void)hipFuncSetAttribute(rs == 8 ? (const void*)k2d_relax_f64 : (const void*)k2d_relax_f32,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->lds2);
  }
  *out = h;
  return DPM_OK;
}

// ---------------------------------------------------- 2D: the wide path

static Run2D run2d_args(dpm_tissue* h, int nsteps, float dt);

// Automatic choice: the wide form for a tissue beyond one workgroup, and for
// a few large tissues, which one workgroup each would keep on a few compute
// units (the 400-cell tissue of the reference's test2D.py: 735 against 1,967
// us per step); many tissues, or small ones, one workgroup per tissue (all
// steps of a call in one launch).
constexpr int kWideMinCells = 128;
constexpr int kWideMaxTissues = 64;
static bool use_wide(const dpm_tissue* h) {
  if (h->rs != 4) return false;
  if (h->wide_mode != 0) return h->wide_mode == 2;
  return !h->fits_wg || (h->max_cells >= kWideMinCells && h->T <= kWideMaxTissues);
}

static int ensure_wide(dpm_tissue* h) {
  if (h->xw) return DPM_OK;
  const int C = h->C, T = h->T;
  int rc = 0;
  rc |= h->alloc(&h->xw, h->V);
  rc |= h->alloc(&h->yw, h->V);
  rc |= h->alloc(&h->cellq, 8 * (size_t)C);
  rc |= h->alloc(&h->ncand_w, C);
  rc |= h->alloc(&h->cand_w, (size_t)C * DPM2D_MAXCAND);
  rc |= h->alloc(&h->cell_tissue2, C);
  rc |= h->alloc(&h->vbl, (size_t)C + T);
  rc |= h->alloc(&h->over_w, ((size_t)h->V + 31) / 32);
  rc |= h->alloc(&h->vcell_w, (size_t)h->V);
  if (rc) return DPM_ENOMEM;
  std::vector<int> ct(C), vbl((size_t)C + T), vc((size_t)h->V);
  for (int t = 0; t < T; t++) {
    const int c0 = h->tissue_begin_h[t], c1 = h->tissue_begin_h[t + 1];
    for (int c = c0; c < c1; c++) ct[c] = t;
    for (int c = c0; c <= c1; c++) vbl[(size_t)c + t] = h->cvb_h[c] - h->cvb_h[c0];
  }
  for (int c = 0; c < C; c++)
    for (int v = h->cvb_h[c]; v < h->cvb_h[c + 1]; v++) vc[v] = c;
  TRY(upload(h->cell_tissue2, ct.data(), C, h->stream));
  TRY(upload(h->vbl, vbl.data(), vbl.size(), h->stream));
  TRY(upload(h->vcell_w, vc.data(), vc.size(), h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  return DPM_OK;
}

static Wide2D wide_args(dpm_tissue* h, int in, float dt, int flags) {
  Wide2D a{};
  a.g = run2d_args(h, 0, 0.0f).g;
  float* X[2] = {(float*)h->x2, h->xw};
  float* Y[2] = {(float*)h->y2, h->yw};
  a.xin = X[in];
  a.yin = Y[in];
  a.xout = X[1 - in];
  a.yout = Y[1 - in];
  a.fx = (float*)h->fx2;
  a.fy = (float*)h->fy2;
  a.cell_tissue = h->cell_tissue2;
  a.vbl = h->vbl;
  a.cellq = h->cellq;
  a.ncand = h->ncand_w;
  a.cand = h->cand_w;
  a.C = h->C;
  a.V = h->V;
  a.dt = dt;
  a.flags = flags;
  a.counters = h->counters;
  a.over_bits = h->over_w;
  a.vcell = h->vcell_w;
  return a;
}

static int launch_wide(dpm_tissue* h, const Wide2D& a) {
  const int bc = (h->C + 255) / 256, bv = (h->V + 255) / 256, bw = (h->C + 3) / 4;
  hipLaunchKernelGGL(k2d_wide_cells, dim3(bc), dim3(256), 0, h->stream, a);
  hipLaunchKernelGGL(k2d_wide_cand, dim3(bw), dim3(256), 0, h->stream, a);  // a wave per cell
  if (a.g.mask & M2_REPULSION) {  // (vertex, candidate slot) threads
    const long long n = (long long)h->V * DPM2D_MAXCAND;
    hipLaunchKernelGGL(k2d_wide_over, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, a);
  }
  hipLaunchKernelGGL(k2d_wide_step, dim3(bv), dim3(256), 0, h->stream, a);
  HIPCHECK(hipGetLastError());
  return DPM_OK;
}

// kGraphSteps steps from buffer 0 (ending on buffer 0), captured once per dt.
static int build_wide_graph(dpm_tissue* h, float dt) {
  if (h->wide_graph) {
    (void)hipGraphExecDestroy(h->wide_graph);
    h->wide_graph = nullptr;
  }
  hipGraph_t g = nullptr;
  HIPCHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int rc = DPM_OK;
  for (int s = 0; s < kGraphSteps && !rc; s++) rc = launch_wide(h, wide_args(h, s & 1, dt, W2_MOVE));
  hipError_t e = hipStreamEndCapture(h->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(DPM_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&h->wide_graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    h->wide_graph = nullptr;
    return fail(DPM_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  h->wide_graph_dt = dt;
  return DPM_OK;
}

// nsteps steps; the positions start and end in buffer 0 (x2, y2), the last
// step writes the forces it used.
static int euler2d_wide(dpm_tissue* h, int nsteps, float dt) {
  TRY(ensure_wide(h));
  if (!h->wide_graph || h->wide_graph_dt != dt) TRY(build_wide_graph(h, dt));
  int cur = 0, left = nsteps - 1;
  while (left >= kGraphSteps) {
    HIPCHECK(hipGraphLaunch(h->wide_graph, h->stream));
    left -= kGraphSteps;
  }
  for (; left > 0; left--, cur ^= 1) TRY(launch_wide(h, wide_args(h, cur, dt, W2_MOVE)));
  TRY(launch_wide(h, wide_args(h, cur, dt, W2_MOVE | W2_FORCES)));
  cur ^= 1;
  if (cur == 1) {
    HIPCHECK(hipMemcpyAsync(h->x2, h->xw, (size_t)h->V * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
    HIPCHECK(hipMemcpyAsync(h->y2, h->yw, (size_t)h->V * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
  }
  return DPM_OK;
}

extern "C" int dpm2d_set_step_kernel(dpm_tissue* h, int kind) {
  if (!h || h->dim != 2) return fail(DPM_EINVAL, "not a 2D tissue");
  if (kind < 0 || kind > 2) return fail(DPM_EINVAL, "kind must be 0 (automatic), 1 (one workgroup) or 2 (wide)");
  if (kind == 1 && !h->fits_wg) return fail(DPM_EINVAL, "a tissue does not fit one workgroup");
  if (kind == 2 && h->rs != 4) return fail(DPM_EINVAL, "the wide 2D step takes float tissues");
  h->wide_mode = kind;
  return DPM_OK;
}

extern "C" int dpm2d_create(const dpm2d_desc* d, dpm_tissue** out) { return create2d(d, out, 4); }
extern "C" int dpm2d_create_f64(const dpm2d_desc* d, dpm_tissue** out) { return create2d(d, out, 8); }

extern "C" int dpm2d_set_cell_params(dpm_tissue* h, const float* Ka, const float* Kl,
                                     const float* Kb, const float* a0, const float* l0,
                                     const float* r0) {
  if (!h || h->dim != 2) return fail(DPM_EINVAL, "not a 2D tissue");
  const int C = h->C;
  hipStream_t s = h->stream;
  if (Ka) TRY(upload(h->Ka, Ka, C, s));
  if (Kl) TRY(upload(h->Kl, Kl, C, s));
  if (Kb) TRY(upload(h->Kb, Kb, C, s));
  if (a0) {
    TRY(upload(h->a0, a0, C, s));
    h->a0_h.assign(a0, a0 + C);
    std::vector<double> sa(h->T, 0.0);  // sum of rest areas per tissue
    for (int t = 0; t < h->T; t++)
      for (int c = h->tissue_begin_h[t]; c < h->tissue_begin_h[t + 1]; c++) sa[t] += a0[c];
    TRY(upload(h->sum_a0, sa.data(), h->T, s));
    HIPCHECK(hipStreamSynchronize(s));
  }
  if (l0) TRY(upload(h->l0_2, l0, C, s));
  if (r0) TRY(upload(h->r0, r0, C, s));
  HIPCHECK(hipStreamSynchronize(s));
  return DPM_OK;
}

template <typename In>
static int upload_real(dpm_tissue* h, void* dst, const In* src, size_t n) {
  if (h->rs == 8) {
    std::vector<double> v(src, src + n);
    TRY(upload((double*)dst, v.data(), n, h->stream));
  } else {
    std::vector<float> v(n);
    for (size_t i = 0; i < n; i++) v[i] = (float)src[i];
    TRY(upload((float*)dst, v.data(), n, h->stream));
  }
  HIPCHECK(hipStreamSynchronize(h->stream));
  return DPM_OK;
}

extern "C" int dpm2d_set_tissue_params(dpm_tissue* h, const float* L, const int* pbc,
                                       const float* Kre, const float* Kat, unsigned mask) {
  if (!h || h->dim != 2) return fail(DPM_EINVAL, "not a 2D tissue");
  const int T = h->T;
  hipStream_t s = h->stream;
  if (L) {
    for (int t = 0; t < T; t++)
      if (!(L[t] > 0) || !std::isfinite(L[t])) return fail(DPM_EINVAL, "box length must be positive");
    TRY(upload_real(h, h->tLr, L, T));
  }
  if (pbc) TRY(upload(h->tpbc, pbc, T, s));
  if (Kre) TRY(upload(h->tKre, Kre, T, s));
  if (Kat) TRY(upload(h->tKat, Kat, T, s));
  h->mask = mask & DPM2D_ALL_TERMS;
  HIPCHECK(hipStreamSynchronize(s));
  return DPM_OK;
}

template <typename In>
static int set_positions2(dpm_tissue* h, const In* xy, size_t n) {
  if (!h || h->dim != 2 || !xy) return fail(DPM_EINVAL, "bad arguments");
  if (n != (size_t)h->V) return fail(DPM_EINVAL, "position count does not match the tissue");
  std::vector<In> a(n), b(n);
  for (size_t i = 0; i < n; i++) {
    a[i] = xy[2 * i];
    b[i] = xy[2 * i + 1];
  }
  TRY(upload_real(h, h->x2, a.data(), n));
  TRY(upload_real(h, h->y2, b.data(), n));
  return DPM_OK;
}

template <typename Out>
static int download2(dpm_tissue* h, const void* a, const void* b, Out* xy, size_t n) {
  auto get = [&](auto tag) -> int {
    using R = decltype(tag);
    std::vector<R> A(n), B(n);
    HIPCHECK(hipMemcpyAsync(A.data(), a, n * sizeof(R), hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipMemcpyAsync(B.data(), b, n * sizeof(R), hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
    for (size_t i = 0; i < n; i++) {
      xy[2 * i] = (Out)A[i];
      xy[2 * i + 1] = (Out)B[i];
    }
    return DPM_OK;
  };
  return h->rs == 8 ? get(0.0) : get(0.0f);
}

template <typename Out>
static int get2(dpm_tissue* h, bool forces, Out* xy, size_t n) {
  if (!h || h->dim != 2 || !xy) return fail(DPM_EINVAL, "bad arguments");
  if (n != (size_t)h->V) return fail(DPM_EINVAL, "count does not match the tissue");
  return forces ? download2(h, h->fx2, h->fy2, xy, n) : download2(h, h->x2, h->y2, xy, n);
}

extern "C" int dpm2d_set_positions(dpm_tissue* h, const float* xy, size_t n) { return set_positions2(h, xy, n); }
extern "C" int dpm2d_get_positions(dpm_tissue* h, float* xy, size_t n) { return get2(h, false, xy, n); }
extern "C" int dpm2d_get_forces(dpm_tissue* h, float* xy, size_t n) { return get2(h, true, xy, n); }
extern "C" int dpm2d_set_positions_f64(dpm_tissue* h, const double* xy, size_t n) {
  return set_positions2(h, xy, n);
}
extern "C" int dpm2d_get_positions_f64(dpm_tissue* h, double* xy, size_t n) { return get2(h, false, xy, n); }
extern "C" int dpm2d_get_forces_f64(dpm_tissue* h, double* xy, size_t n) { return get2(h, true, xy, n); }

extern "C" int dpm2d_get_box(dpm_tissue* h, double* L) {
  if (!h || h->dim != 2 || !L) return fail(DPM_EINVAL, "bad arguments");
  std::vector<double> out(h->T);
  if (h->rs == 8) {
    HIPCHECK(hipMemcpyAsync(out.data(), h->tLr, h->T * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
  } else {
    std::vector<float> f(h->T);
    HIPCHECK(hipMemcpyAsync(f.data(), h->tLr, h->T * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
    for (int t = 0; t < h->T; t++) out[t] = f[t];
  }
  std::copy(out.begin(), out.end(), L);
  return DPM_OK;
}

template <typename R>
static Geom2D<R> geom2d(dpm_tissue* h) {
  Geom2D<R> g{};
  g.tissue_cell_begin = h->tcb;
  g.cell_vert_begin = h->cvb;
  g.Ka = h->Ka;
  g.Kl = h->Kl;
  g.Kb = h->Kb;
  g.a0 = h->a0;
  g.l0 = h->l0_2;
  g.r0 = h->r0;
  g.tL = (const R*)h->tLr;
  g.tpbc = h->tpbc;
  g.tKre = h->tKre;
  g.tKat = h->tKat;
  g.mask = h->mask;
  return g;
}

static Run2D run2d_args(dpm_tissue* h, int nsteps, float dt) {
  Run2D a{};
  a.g = geom2d<float>(h);
  a.x = (float*)h->x2;
  a.y = (float*)h->y2;
  a.fx = (float*)h->fx2;
  a.fy = (float*)h->fy2;
  a.nsteps = nsteps;
  a.dt = dt;
  a.counters = h->counters;
  a.over_cap = h->over_cap;
  return a;
}

static FireParams fire_params(const dpm_fire_params* p) {
  FireParams f{};
  f.dt0 = p->dt0;
  f.dt_max = p->dt_max;
  f.ftol = p->ftol;
  f.f_inc = p->f_inc;
  f.f_dec = p->f_dec;
  f.alpha0 = p->alpha0;
  f.f_alpha = p->f_alpha;
  f.max_iter = p->max_iter;
  f.n_min = p->n_min;
  return f;
}

extern "C" void dpm_fire_defaults(dpm_fire_params* p) {
  if (!p) return;
  p->dt0 = 0.01;
  p->dt_max = 0.1;
  p->ftol = 1e-10;
  p->f_inc = 1.1;
  p->f_dec = 0.5;
  p->alpha0 = 0.1;
  p->f_alpha = 0.99;
  p->max_iter = 100000;
  p->n_min = 5;
}

extern "C" void dpm_jam_defaults(dpm_jam_params* p) {
  if (!p) return;
  p->phi0 = 0.35;
  p->dphi = 0.02;
  p->phi_tol = 1e-3;
  p->u_jam = 1e-16;
  p->max_stages = 40;
  dpm_fire_defaults(&p->fire);
  p->fire.max_iter = 20000;
}

static int check_fire(const dpm_fire_params* p) {
  if (!p) return fail(DPM_EINVAL, "null FIRE parameters");
  if (!(p->dt0 > 0) || !(p->dt_max >= p->dt0) || !(p->ftol >= 0) || !(p->f_inc >= 1) || !(p->f_dec > 0) ||
      !(p->f_dec < 1) || !(p->alpha0 >= 0) || !(p->alpha0 < 1) || !(p->f_alpha > 0) || !(p->f_alpha <= 1) ||
      p->max_iter < 0 || p->n_min < 0)
    return fail(DPM_EINVAL, "invalid FIRE parameters");
  return DPM_OK;
}

template <typename Edit>
static int launch_relax(dpm_tissue* h, int mode, int threads, Edit edit) {
  auto go = [&](auto tag) -> int {
    using R = decltype(tag);
    Relax2D<R> a{};
    a.g = geom2d<R>(h);
    a.x = (R*)h->x2;
    a.y = (R*)h->y2;
    a.vx = (R*)h->vx2;
    a.vy = (R*)h->vy2;
    a.fx = (R*)h->fx2;
    a.fy = (R*)h->fy2;
    a.sx = (R*)h->sx2;
    a.sy = (R*)h->sy2;
    a.tLout = (R*)h->tLr;
    a.res = h->res2;
    a.sum_a0 = h->sum_a0;
    a.mode = mode;
    dpm_fire_params fd;
    dpm_fire_defaults(&fd);
    a.fire = fire_params(&fd);
    edit(a);
    void (*k)(Relax2D<R>) = nullptr;
    if constexpr (sizeof(R) == 8) k = k2d_relax_f64; else k = k2d_relax_f32;
    hipLaunchKernelGGL(k, dim3(h->T), dim3(threads > 0 ? threads : h->NTr), h->lds2, h->stream, a);
    HIPCHECK(hipGetLastError());
    return DPM_OK;
  };
  return h->rs == 8 ? go(0.0) : go(0.0f);
}

static int ensure_relax_buffers(dpm_tissue* h) {
  if (!h->fits_wg)
    return fail(DPM_ERANGE, "FIRE, jamming and energies need 2D tissues that fit one workgroup's LDS");
  if (h->vx2) return DPM_OK;
  int rc = h->alloc_real(&h->vx2, h->V);
  rc |= h->alloc_real(&h->vy2, h->V);
  rc |= h->alloc_real(&h->sx2, h->V);
  rc |= h->alloc_real(&h->sy2, h->V);
  return rc ? DPM_ENOMEM : DPM_OK;
}

static int read_stats(dpm_tissue* h, dpm_relax_stats* st) {
  if (st)
    HIPCHECK(hipMemcpyAsync(st, h->res2, sizeof(double) * h->T * RS_COUNT, hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  return DPM_OK;
}

static int euler2d(dpm_tissue* h, int nsteps, float dt) {
  if (nsteps <= 0) return fail(DPM_EINVAL, "nsteps must be positive");
  if (h->rs == 8) {
    TRY(ensure_relax_buffers(h));
    return launch_relax(h, RELAX_EULER, 0, [&](auto& a) {
      a.nsteps = nsteps;
      a.dt = dt;
    });
  }
  if (use_wide(h)) return euler2d_wide(h, nsteps, dt);
  Run2D a = run2d_args(h, nsteps, dt);
  hipLaunchKernelGGL(k2d_run, dim3(h->T), dim3(h->NT2), h->lds_run, h->stream, a);
  HIPCHECK(hipGetLastError());
  return DPM_OK;
}

extern "C" int dpm2d_set_relax_threads(dpm_tissue* h, int threads) {
  if (!h || h->dim != 2) return fail(DPM_EINVAL, "not a 2D tissue");
  if (threads == 0) threads = std::min(1024, std::max(64, ((h->V / std::max(1, h->T)) / 4 + 63) / 64 * 64));
  if (threads < 64 || threads > 1024 || threads % 64) return fail(DPM_EINVAL, "threads must be 64..1024, a multiple of 64");
  h->NTr = threads;
  return DPM_OK;
}

extern "C" int dpm2d_relax_threads(dpm_tissue* h) { return (h && h->dim == 2) ? h->NTr : 0; }

extern "C" int dpm2d_fire(dpm_tissue* h, const dpm_fire_params* p, dpm_relax_stats* stats) {
  if (!h || h->dim != 2) return fail(DPM_EINVAL, "not a 2D tissue");
  TRY(check_fire(p));
  TRY(ensure_relax_buffers(h));
  TRY(launch_relax(h, RELAX_FIRE, 0, [&](auto& a) { a.fire = fire_params(p); }));
  return read_stats(h, stats);
}

extern "C" int dpm2d_energy(dpm_tissue* h, dpm_relax_stats* stats) {
  if (!h || h->dim != 2) return fail(DPM_EINVAL, "not a 2D tissue");
  TRY(ensure_relax_buffers(h));
  TRY(launch_relax(h, RELAX_ENERGY, 0, [](auto&) {}));
  return read_stats(h, stats);
}

static int check_jam(dpm_tissue* h, const dpm_jam_params* p) {
  if (!h || h->dim != 2) return fail(DPM_EINVAL, "not a 2D tissue");
  if (!p) return fail(DPM_EINVAL, "null jamming parameters");
  TRY(check_fire(&p->fire));
  if (!(p->phi0 > 0) || !(p->phi0 < 1) || !(p->dphi > 0) || !(p->phi_tol > 0) || !(p->u_jam >= 0) ||
      p->max_stages <= 0)
    return fail(DPM_EINVAL, "invalid jamming parameters");
  if (!(h->mask & DPM2D_VREPULSION))
    return fail(DPM_EINVAL, "jamming classifies by the vertex-repulsion energy: enable DPM2D_VREPULSION");
  return ensure_relax_buffers(h);
}

template <typename A>
static void jam_args(A& a, const dpm_jam_params* p, int start, int stages) {
  a.fire = fire_params(&p->fire);
  a.jam.phi0 = p->phi0;
  a.jam.dphi = p->dphi;
  a.jam.phi_tol = p->phi_tol;
  a.jam.u_jam = p->u_jam;
  a.jam.max_stages = p->max_stages;
  a.jam_start = start;
  a.jam_stages = stages;
}

// The whole jamming protocol is one launch, captured into a hipGraph per
// handle and parameter set and replayed by later calls with the same
// parameters.
extern "C" int dpm2d_jam(dpm_tissue* h, const dpm_jam_params* p, dpm_relax_stats* stats) {
  TRY(check_jam(h, p));
  std::vector<unsigned char> key(sizeof(dpm_jam_params) + 2 * sizeof(int));
  std::memcpy(key.data(), p, sizeof(dpm_jam_params));
  const int extra[2] = {h->NTr, (int)h->mask};
  std::memcpy(key.data() + sizeof(dpm_jam_params), extra, sizeof(extra));
  if (!h->jam_graph || key != h->jam_key) {
    if (h->jam_graph) {
      (void)hipGraphExecDestroy(h->jam_graph);
      h->jam_graph = nullptr;
    }
    hipGraph_t g = nullptr;
    HIPCHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const int rc = launch_relax(h, RELAX_JAM, 0, [&](auto& a) { jam_args(a, p, 1, p->max_stages); });
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(DPM_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&h->jam_graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      h->jam_graph = nullptr;
      return fail(DPM_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    h->jam_key = key;
  }
  HIPCHECK(hipGraphLaunch(h->jam_graph, h->stream));
  return read_stats(h, stats);
}

extern "C" int dpm2d_jam_step(dpm_tissue* h, const dpm_jam_params* p, int start, int stages,
                              dpm_relax_stats* stats) {
  TRY(check_jam(h, p));
  if (stages <= 0) return fail(DPM_EINVAL, "stages must be positive");
  TRY(launch_relax(h, RELAX_JAM, 0, [&](auto& a) { jam_args(a, p, start ? 1 : 0, stages); }));
  return read_stats(h, stats);
}

// Workgroup-size autotuner of the relaxation kernel: times `iters` FIRE
// iterations (ftol 0, so every packing runs all of them) for each candidate
// size from the current positions, restores the positions after each trial
// and keeps the fastest.  times_ms (may be NULL) receives the five timings
// of 64, 128, 256, 512, 1024 threads (-1 where not run).
extern "C" int dpm2d_autotune(dpm_tissue* h, int iters, int* best, float* times_ms) {
  if (!h || h->dim != 2) return fail(DPM_EINVAL, "not a 2D tissue");
  if (iters <= 0) return fail(DPM_EINVAL, "iters must be positive");
  TRY(ensure_relax_buffers(h));
  const size_t bytes = (size_t)h->V * h->rs;
  void *bx = nullptr, *by = nullptr;
  HIPCHECK(hipMalloc(&bx, bytes));
  if (hipMalloc(&by, bytes) != hipSuccess) {
    (void)hipFree(bx);
    return fail(DPM_ENOMEM, "autotune buffers");
  }
  hipStream_t s = h->stream;
  int rc = DPM_OK;
  if (hipMemcpyAsync(bx, h->x2, bytes, hipMemcpyDeviceToDevice, s) ||
      hipMemcpyAsync(by, h->y2, bytes, hipMemcpyDeviceToDevice, s))
    rc = fail(DPM_EDEVICE, "autotune copy");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && (hipEventCreate(&e0) || hipEventCreate(&e1))) rc = fail(DPM_EDEVICE, "autotune events");
  const int cand[5] = {64, 128, 256, 512, 1024};
  float bestt = -1.0f;
  int bestn = h->NTr;
  dpm_fire_params fp;
  dpm_fire_defaults(&fp);
  fp.ftol = 0.0;
  fp.max_iter = iters;
  for (int i = 0; i < 5 && !rc; i++) {
    float ms = -1.0f;
    (void)hipEventRecord(e0, s);
    rc = launch_relax(h, RELAX_FIRE, cand[i], [&](auto& a) { a.fire = fire_params(&fp); });
    (void)hipEventRecord(e1, s);
    if (!rc && hipEventSynchronize(e1) == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
    if (times_ms) times_ms[i] = ms;
    if (!rc && ms > 0 && (bestt < 0 || ms < bestt)) {
      bestt = ms;
      bestn = cand[i];
    }
    if (hipMemcpyAsync(h->x2, bx, bytes, hipMemcpyDeviceToDevice, s) ||
        hipMemcpyAsync(h->y2, by, bytes, hipMemcpyDeviceToDevice, s))
      rc = fail(DPM_EDEVICE, "autotune restore");
  }
  (void)hipStreamSynchronize(s);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(bx);
  (void)hipFree(by);
  if (rc) return rc;
  h->NTr = bestn;
  if (best) *best = bestn;
  return DPM_OK;
}

extern "C" int dpm2d_probe(dpm_tissue* h, float* forces, int* overlap, int* attraction) {
  if (!h || h->dim != 2) return fail(DPM_EINVAL, "not a 2D tissue");
  if (h->rs == 8) {
    if (overlap || attraction)
      return fail(DPM_EINVAL, "pair diagnostics are recorded by float tissues only");
    TRY(dpm2d_energy(h, nullptr));
    if (forces) TRY(dpm2d_get_forces(h, forces, h->V));
    return DPM_OK;
  }
  hipStream_t s = h->stream;
  int *dov = nullptr, *dat = nullptr;
  int rc = dev_alloc(&dov, h->V);
  rc |= dev_alloc(&dat, h->V);
  if (!rc && use_wide(h)) rc = ensure_wide(h);
  if (!rc) {
    (void)hipMemsetAsync(dov, 0, h->V * sizeof(int), s);
    (void)hipMemsetAsync(dat, 0, h->V * sizeof(int), s);
    if (use_wide(h)) {
      Wide2D a = wide_args(h, 0, 0.0f, W2_FORCES);
      a.counters = nullptr;
      a.p_overlap = dov;
      a.p_attr = dat;
      rc = launch_wide(h, a);
    } else {
      Run2D a = run2d_args(h, 0, 0.0f);
      a.counters = nullptr;
      a.p_overlap = dov;
      a.p_attr = dat;
      hipLaunchKernelGGL(k2d_run, dim3(h->T), dim3(h->NT2), h->lds_run, s, a);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) rc = fail(DPM_EDEVICE, std::string("k2d_run: ") + hipGetErrorString(e));
    }
  }
  if (!rc && overlap)
    rc = hipMemcpyAsync(overlap, dov, h->V * sizeof(int), hipMemcpyDeviceToHost, s) ? DPM_EDEVICE : 0;
  if (!rc && attraction)
    rc = hipMemcpyAsync(attraction, dat, h->V * sizeof(int), hipMemcpyDeviceToHost, s) ? DPM_EDEVICE : 0;
  if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = fail(DPM_EDEVICE, "stream sync failed");
  if (dov) (void)hipFree(dov);
  if (dat) (void)hipFree(dat);
  if (rc) return rc;
  if (forces) TRY(dpm2d_get_forces(h, forces, h->V));
  return DPM_OK;
}

extern "C" int dpm2d_probe_attraction(dpm_tissue* h, unsigned* keys) {
  if (!h || h->dim != 2) return fail(DPM_EINVAL, "not a 2D tissue");
  if (h->rs == 8) return fail(DPM_EINVAL, "pair diagnostics are recorded by float tissues only");
  if (!keys) return DPM_OK;
  hipStream_t s = h->stream;
  unsigned* dk = nullptr;
  int rc = dev_alloc(&dk, h->V);
  if (!rc && use_wide(h)) rc = ensure_wide(h);
  if (!rc) {
    (void)hipMemsetAsync(dk, 0, h->V * sizeof(unsigned), s);
    if (use_wide(h)) {
      Wide2D a = wide_args(h, 0, 0.0f, 0);
      a.counters = nullptr;
      a.p_attr_key = dk;
      rc = launch_wide(h, a);
    } else {
      Run2D a = run2d_args(h, 0, 0.0f);
      a.counters = nullptr;
      a.p_attr_key = dk;
      hipLaunchKernelGGL(k2d_run, dim3(h->T), dim3(h->NT2), h->lds_run, s, a);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) rc = fail(DPM_EDEVICE, std::string("k2d_run: ") + hipGetErrorString(e));
    }
  }
  if (!rc)
    rc = hipMemcpyAsync(keys, dk, h->V * sizeof(unsigned), hipMemcpyDeviceToHost, s) ? DPM_EDEVICE : 0;
  if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = fail(DPM_EDEVICE, "stream sync failed");
  if (dk) (void)hipFree(dk);
  return rc;
}

// ============================================================= dispersal ==

extern "C" int dpm_disperse(const float* r, int n, float L, float* X, float* Y, int max_count, int* result) {
  if (!r || !X || !Y || n < 1) return fail(DPM_EINVAL, "dpm_disperse: bad arguments");
  if (n > 65535) return fail(DPM_ERANGE, "dpm_disperse: more than 65535 centres");
  if (!(L > 0.0f) || !std::isfinite(L)) return fail(DPM_EINVAL, "dpm_disperse: box length must be positive");
  if (max_count < 0) return fail(DPM_EINVAL, "dpm_disperse: max_count must be >= 0");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(DPM_EDEVICE, "no HIP device is visible");
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  float rmax = 0.0f;
  for (int i = 0; i < n; i++) {
    if (!(r[i] >= 0.0f) || !std::isfinite(r[i])) return fail(DPM_EINVAL, "dpm_disperse: radii must be finite and >= 0");
    if (!std::isfinite(X[i]) || !std::isfinite(Y[i])) return fail(DPM_ENONFINITE, "dpm_disperse: non-finite centre");
    rmax = std::max(rmax, r[i]);
  }
  // bins no narrower than the largest contact distance 2 rmax (with slack)
  int nb = rmax > 0.0f ? (int)std::floor(L / (2.0f * rmax * 1.001f + 1e-6f)) : 0;
  if (nb < 3 || nb > 256) nb = nb > 256 ? 256 : 0;
  const int threads = std::min(kDispThreads, std::max(64, (n + 63) / 64 * 64));
  size_t lds = disp_layout(nullptr, n, nb, threads).total;
  const int maxlds = max_lds_per_block(dev);
  if ((int)lds > maxlds && nb > 0) {
    nb = 0;
    lds = disp_layout(nullptr, n, nb, threads).total;
  }
  if ((int)lds > maxlds) return fail(DPM_ERANGE, "dpm_disperse: the centres do not fit the LDS of one workgroup");
  const int term_cap = n * kDispMaxCand;
  hipStream_t s = nullptr;
  HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  float *dr = nullptr, *dx = nullptr, *dy = nullptr, *dt = nullptr;
  int* dout = nullptr;
  int rc = dev_alloc(&dr, n);
  rc |= dev_alloc(&dx, n);
  rc |= dev_alloc(&dy, n);
  rc |= dev_alloc(&dout, 2);
  rc |= dev_alloc(&dt, (size_t)term_cap);
  int out[2] = {0, 0};
  if (!rc) rc = upload(dr, r, n, s) | upload(dx, X, n, s) | upload(dy, Y, n, s);
  if (!rc && hipFuncSetAttribute((const void*)k_disperse2d, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
                 hipSuccess)
    rc = fail(DPM_EDEVICE, "hipFuncSetAttribute(k_disperse2d) failed");
  if (!rc) {
    Disperse2DArgs a{dr, dx, dy, n, L, max_count, nb, dout, dt, term_cap};
    hipLaunchKernelGGL(k_disperse2d, dim3(1), dim3(threads), lds, s, a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = fail(DPM_EDEVICE, std::string("k_disperse2d: ") + hipGetErrorString(e));
  }
  if (!rc && (hipMemcpyAsync(X, dx, n * sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess ||
              hipMemcpyAsync(Y, dy, n * sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess ||
              hipMemcpyAsync(out, dout, 2 * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess))
    rc = fail(DPM_EDEVICE, "dpm_disperse: copy back failed");
  if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = fail(DPM_EDEVICE, "dpm_disperse: stream sync failed");
  for (void* p : {(void*)dr, (void*)dx, (void*)dy, (void*)dout, (void*)dt})
    if (p) (void)hipFree(p);
  (void)hipStreamDestroy(s);
  if (rc) return rc;
  if (result) {
    result[0] = out[0];
    result[1] = out[1];
  }
  return DPM_OK;
}

// ================================================================ common ==

extern "C" int dpm_euler_async(dpm_tissue* h, int nsteps, float dt) {
  if (!h) return fail(DPM_EINVAL, "null handle");
  if (h->dim == 3) return euler3d(h, nsteps, dt);
  return euler2d(h, nsteps, dt);
}

extern "C" int dpm_euler(dpm_tissue* h, int nsteps, float dt) {
  TRY(dpm_euler_async(h, nsteps, dt));
  if (h->dim == 3) {
    const size_t n = (size_t)h->C * h->nv;
    const int c = h->cur;
    if (h->rs == 8)
      return check_finite(h, (const double*)h->x[c], (const double*)h->y[c], (const double*)h->z[c], n);
    return check_finite(h, (const float*)h->x[c], (const float*)h->y[c], (const float*)h->z[c], n);
  }
  // The reference's 2D path performs no result check (Tissue2D.cpp:231-240).
  HIPCHECK(hipStreamSynchronize(h->stream));
  return DPM_OK;
}
