// 2D DPM timestep kernels for gfx950.
//
// A 2D tissue is small enough to live entirely in one workgroup's LDS, so a
// whole CLEulerUpdate call (nsteps steps) — or a whole FIRE minimisation —
// is ONE launch with one workgroup per tissue: a batch of independent
// tissues of uneven size runs as one grid.  Steps are separated by
// workgroup barriers only; nothing goes back to HBM until the end.
//
// Reference semantics followed (Cell2D_kernel.cl):
//   AreaForceUpdates       :13-42   (area summed in double per cell)
//   PerimeterForceUpdates  :92-119
//   BendingForceUpdates    :44-90
//   AttractionForceUpdate  :222-268 (vertex pairs closer than l0 of own cell)
//   RepulsionForceUpdate   :121-220 (crossing parity with its PBC quirk)
//   EulerUpdate            :270-281
// Pair predicates (d < l0, crossing parity) are evaluated with the
// reference's operation order in IEEE float; cells are skipped only where
// the predicate is provably false for every vertex pair (guards below).
#include "dpm_device.h"

using namespace dpm;

#define DPM2D_MAXCAND 48

enum : unsigned {
  M2_AREA = 1u,
  M2_PERIMETER = 2u,
  M2_BENDING = 4u,
  M2_ATTRACTION = 8u,
  M2_REPULSION = 16u,
};

struct Run2D {
  float* x;  // [nverts] packed cell after cell, all tissues
  float* y;
  float* fx;  // forces of the last evaluation (written at the end)
  float* fy;
  const int* tissue_cell_begin;  // [T+1]
  const int* cell_vert_begin;    // [C+1] (global vertex offsets)
  const float* Ka;
  const float* Kl;
  const float* Kb;
  const float* a0;
  const float* l0;
  const float* r0;
  const float* tL;
  const int* tpbc;
  const float* tKre;
  const float* tKat;
  unsigned mask;
  int nsteps;  // Euler steps (0 = evaluate forces only)
  float dt;
  unsigned long long* counters;  // [4]
  int* p_overlap;                // diagnostics, per vertex (nullptr normally)
  int* p_attr;
};

// LDS per tissue: x,y [nvt] float | vcell [nvt] short |
// per cell: vb [nc+1] int, strain, comx, comy, ax0, ax1, ay0, ay1 [nc] float,
// ncand [nc] int, cand [nc*MAXCAND] short | scratch.  Forces stay in
// registers (up to DPM2D_VPT vertices per thread).
#define DPM2D_VPT 16
__host__ __device__ inline size_t lds2d_cells_offset(int nvt) {
  size_t b = sizeof(float) * 2u * (size_t)nvt + sizeof(short) * (size_t)nvt;
  return (b + 3u) & ~(size_t)3u;
}
__host__ __device__ inline size_t lds2d_bytes(int nvt, int nc) {
  size_t b = lds2d_cells_offset(nvt);
  b += sizeof(int) * (size_t)(nc + 1) + sizeof(float) * 7u * (size_t)nc + sizeof(int) * (size_t)nc;
  b += sizeof(short) * (size_t)nc * DPM2D_MAXCAND;
  b = (b + 7u) & ~(size_t)7u;
  return b + sizeof(double) * 32u;
}

struct Lds2D {
  float *x, *y;
  short* vcell;
  int* vb;
  float *strain, *comx, *comy, *ax0, *ax1, *ay0, *ay1;
  int* ncand;
  short* cand;
  double* red;
};

__device__ __forceinline__ Lds2D carve2d(unsigned char* smem, int nvt, int nc) {
  Lds2D l;
  float* f = reinterpret_cast<float*>(smem);
  l.x = f;
  l.y = f + nvt;
  l.vcell = reinterpret_cast<short*>(f + 2 * nvt);
  l.vb = reinterpret_cast<int*>(smem + lds2d_cells_offset(nvt));
  float* cf = reinterpret_cast<float*>(l.vb + nc + 1);
  l.strain = cf;
  l.comx = cf + nc;
  l.comy = cf + 2 * nc;
  l.ax0 = cf + 3 * nc;
  l.ax1 = cf + 4 * nc;
  l.ay0 = cf + 5 * nc;
  l.ay1 = cf + 6 * nc;
  l.ncand = reinterpret_cast<int*>(cf + 7 * nc);
  l.cand = reinterpret_cast<short*>(l.ncand + nc);
  size_t off = lds2d_cells_offset(nvt);
  off += sizeof(int) * (size_t)(nc + 1) + sizeof(float) * 7u * (size_t)nc + sizeof(int) * (size_t)nc;
  off += sizeof(short) * (size_t)nc * DPM2D_MAXCAND;
  off = (off + 7u) & ~(size_t)7u;
  l.red = reinterpret_cast<double*>(smem + off);
  return l;
}

struct Tissue2DCtx {
  int c0, nc, v0, nvt;  // first cell, cells, first vertex, vertices
  float L;
  int pbc;
  float Kre, Kat;
};

// Per-cell quantities of the current positions: area strain (double area),
// centre of mass, axis-aligned bounds.  One thread per cell.
__device__ void cell_pass(const Run2D& a, const Tissue2DCtx& T, const Lds2D& l) {
  for (int c = threadIdx.x; c < T.nc; c += blockDim.x) {
    const int b = l.vb[c], e = l.vb[c + 1], n = e - b;
    double area = 0.0, sx = 0.0, sy = 0.0;
    float x0 = l.x[b], x1 = l.x[b], y0 = l.y[b], y1 = l.y[b];
    for (int j = 0; j < n; j++) {
      const int k = b + j, km = (j == 0) ? (e - 1) : (k - 1);
      const double xs = (double)l.x[km] + (double)l.x[k];
      const double yd = (double)l.y[km] - (double)l.y[k];
      area += xs * yd;
      sx += (double)l.x[k];
      sy += (double)l.y[k];
      x0 = fminf(x0, l.x[k]);
      x1 = fmaxf(x1, l.x[k]);
      y0 = fminf(y0, l.y[k]);
      y1 = fmaxf(y1, l.y[k]);
    }
    area = 0.5 * fabs(area);
    const int gc = T.c0 + c;
    l.strain[c] = (float)(area / (double)a.a0[gc] - 1.0);
    l.comx[c] = (float)(sx / n);
    l.comy[c] = (float)(sy / n);
    l.ax0[c] = x0;
    l.ax1[c] = x1;
    l.ay0[c] = y0;
    l.ay1[c] = y1;
  }
}

// Cell-level candidate lists: cell cj is a candidate of ci unless every
// vertex pair is provably outside both pair predicates.
//  - attraction (d < l0[ci] under the per-pair minimum image): per axis the
//    minimum-image gap between the two boxes exceeds l0 and the boxes are
//    small against L (then no other image can come closer);
//  - repulsion (crossing parity of raw differences, transformed only where a
//    difference exceeds L): the raw boxes are disjoint and no raw
//    difference can exceed L (then the parity of a point outside cj's box
//    is even).
__device__ bool maybe_pair(const Run2D& a, const Tissue2DCtx& T, const Lds2D& l, int ci, int cj) {
  const float L = T.L;
  const float tiny = 1e-5f;
  // raw-coordinate spans of the pair of cells
  const float spanx = fmaxf(l.ax1[ci] - l.ax0[cj], l.ax1[cj] - l.ax0[ci]);
  const float spany = fmaxf(l.ay1[ci] - l.ay0[cj], l.ay1[cj] - l.ay0[ci]);
  if (a.mask & M2_REPULSION) {
    const bool raw_guard = !T.pbc || (spanx < L * 0.999f && spany < L * 0.999f);
    if (!raw_guard) return true;
    const float m = tiny * (1.0f + fmaxf(fabsf(l.ax1[cj]), fabsf(l.ax0[cj])) +
                            fmaxf(fabsf(l.ay1[cj]), fabsf(l.ay0[cj])));
    const bool disjoint = l.ax1[ci] < l.ax0[cj] - m || l.ax0[ci] > l.ax1[cj] + m ||
                          l.ay1[ci] < l.ay0[cj] - m || l.ay0[ci] > l.ay1[cj] + m;
    if (!disjoint) return true;
  }
  if (a.mask & M2_ATTRACTION) {
    const float l0i = a.l0[T.c0 + ci];
    const float hx = 0.5f * ((l.ax1[ci] - l.ax0[ci]) + (l.ax1[cj] - l.ax0[cj]));
    const float hy = 0.5f * ((l.ay1[ci] - l.ay0[ci]) + (l.ay1[cj] - l.ay0[cj]));
    float dx = 0.5f * (l.ax0[cj] + l.ax1[cj]) - 0.5f * (l.ax0[ci] + l.ax1[ci]);
    float dy = 0.5f * (l.ay0[cj] + l.ay1[cj]) - 0.5f * (l.ay0[ci] + l.ay1[ci]);
    if (T.pbc) {
      if (!(hx + l0i < 0.45f * L && hy + l0i < 0.45f * L)) return true;
      dx -= L * roundf(dx / L);
      dy -= L * roundf(dy / L);
    }
    const float reach = l0i * 1.0001f + tiny * (1.0f + L);
    const bool far = fabsf(dx) - hx > reach || fabsf(dy) - hy > reach;
    if (!far) return true;
  }
  return false;
}

__device__ void cand_pass(const Run2D& a, const Tissue2DCtx& T, const Lds2D& l) {
  for (int ci = threadIdx.x; ci < T.nc; ci += blockDim.x) {
    int n = 0;
    for (int cj = 0; cj < T.nc; cj++) {
      if (cj == ci) continue;
      if (maybe_pair(a, T, l, ci, cj)) {
        if (n < DPM2D_MAXCAND) l.cand[ci * DPM2D_MAXCAND + n] = (short)cj;
        n++;
      }
    }
    l.ncand[ci] = n;  // > MAXCAND: the vertex loops scan every cell
  }
}

// Crossing parity of point p against cell cj, Cell2D_kernel.cl:171-196.
__device__ __forceinline__ bool crossing_odd(const Lds2D& l, int cj, float px, float py, int pbc, float L) {
  const int b = l.vb[cj], e = l.vb[cj + 1], n = e - b;
  bool odd = false;
  for (int i = 0, j = n - 1; i < n; j = i++) {
    float dix = px - l.x[b + i], diy = py - l.y[b + i];
    float djx = px - l.x[b + j], djy = py - l.y[b + j];
    if (pbc) {
      if (fabsf(dix) > L || fabsf(djx) > L) {
        dix -= L * floorf(dix / L);
        djx -= L * floorf(djx / L);
      }
      if (fabsf(diy) > L || fabsf(djy) > L) {
        diy -= L * roundf(diy / L);
        djy -= L * roundf(djy / L);
      }
    }
    if ((diy > 0.0f) != (djy > 0.0f)) {
      const float tt = (djx - dix) * (0.0f - diy) / (djy - diy) + dix;
      if (0.0f < tt) odd = !odd;
    }
  }
  return odd;
}

// All force terms on vertex k (tissue-local index) of cell ci.
__device__ __noinline__ void force_vertex(const Run2D& a, const Tissue2DCtx& T, const Lds2D& l, int k,
                             float& Fx, float& Fy, int* n_over, int* n_attr) {
  const int ci = (int)l.vcell[k];
  const int gc = T.c0 + ci;
  const int b = l.vb[ci], e = l.vb[ci + 1], n = e - b, vi = k - b;
  const int im1 = (vi == 0) ? e - 1 : k - 1;
  const int ip1 = (vi == n - 1) ? b : k + 1;
  const float px = l.x[k], py = l.y[k];
  const float L = T.L;
  float fx = 0.0f, fy = 0.0f;
  const unsigned mask = a.mask;
  if (mask & M2_AREA) {
    const float Ka = a.Ka[gc], a0 = a.a0[gc];
    const double pre = (double)(Ka / sqrtf(a0)) * 0.5 * (double)l.strain[ci];
    fx = (float)((double)fx + pre * (double)(l.y[im1] - l.y[ip1]));
    fy = (float)((double)fy + pre * (double)(l.x[im1] - l.x[ip1]));
  }
  if (mask & M2_PERIMETER) {
    const float l0 = a.l0[gc];
    const float lvx = l.x[ip1] - px, lvy = l.y[ip1] - py;
    const float lvmx = px - l.x[im1], lvmy = py - l.y[im1];
    const float len = sqrtf(lvx * lvx + lvy * lvy);
    const float lenm = sqrtf(lvmx * lvmx + lvmy * lvmy);
    const float dli = (float)((double)(len / l0) - 1.0);
    const float dlim1 = (float)((double)(lenm / l0) - 1.0);
    const float pref = a.Kl[gc] * sqrtf(a.a0[gc] / l0);
    fx += pref * (dli * (lvx / len) - dlim1 * (lvmx / lenm));
    fy += pref * (dli * (lvy / len) - dlim1 * (lvmy / lenm));
  }
  if (mask & M2_BENDING) {
    const int ip2 = (vi == n - 1) ? b + 1 : ((vi == n - 2) ? b : k + 2);
    const int im2 = (vi == 0) ? e - 2 : ((vi == 1) ? e - 1 : k - 2);
    const float lvx = l.x[ip1] - px, lvy = l.y[ip1] - py;
    const float lvxm = px - l.x[im1], lvym = py - l.y[im1];
    const float six = lvx - lvxm, siy = lvy - lvym;
    const float sixp = (l.x[ip2] - l.x[ip1]) - lvx, siyp = (l.y[ip2] - l.y[ip1]) - lvy;
    const float sixm = lvxm - (l.x[im1] - l.x[im2]), siym = lvym - (l.y[im1] - l.y[im2]);
    const double Kb = (double)a.Kb[gc];
    fx = (float)((double)fx + Kb * (2.0 * (double)six - (double)sixm - (double)sixp));
    fy = (float)((double)fy + Kb * (2.0 * (double)siy - (double)siym - (double)siyp));
  }
  const int nc = l.ncand[ci];
  const bool scan_all = nc > DPM2D_MAXCAND;
  const int ncl = scan_all ? T.nc : nc;
  if (mask & M2_ATTRACTION) {
    const float l0 = a.l0[gc];
    const float kat = T.Kat / (float)n;
    const float half_l = 0.49f * L;
    const float lim = l0 * l0 * 1.00001f;
    int cnt = 0;
    for (int q = 0; q < ncl; q++) {
      const int cj = scan_all ? q : (int)l.cand[ci * DPM2D_MAXCAND + q];
      if (cj == ci) continue;
      for (int j = l.vb[cj]; j < l.vb[cj + 1]; j++) {
        float rx = l.x[j] - px, ry = l.y[j] - py;
        float nx = 0.0f, ny = 0.0f;
        // minimum image; a component below 0.49 L rounds to image 0, which
        // leaves it unchanged, so the division is skipped there
        if (T.pbc && !(fabsf(rx) < half_l && fabsf(ry) < half_l)) {
          nx = roundf(rx / L);
          ny = roundf(ry / L);
          rx -= L * nx;
          ry -= L * ny;
        }
        const float d2 = rx * rx + ry * ry;
        if (!(d2 < lim)) continue;  // sqrt(d2) < l0 needs d2 < l0^2 (1 + 1e-5)
        const float dist = sqrtf(d2);
        if (dist < l0) {
          // the reference's float decision; the force (kat d/l0 along r)
          // in double on the same image, as a short image vector keeps few
          // digits in float
          const double ddx = ((double)l.x[j] - (double)px) - (double)L * (double)nx;
          const double ddy = ((double)l.y[j] - (double)py) - (double)L * (double)ny;
          const double k = (double)kat / (double)l0;  // ftmp/d = kat/l0
          fx += (float)(k * ddx);
          fy += (float)(k * ddy);
          cnt++;
        }
      }
    }
    *n_attr = cnt;
  }
  if (mask & M2_REPULSION) {
    bool over = false;
    for (int q = 0; q < ncl && !over; q++) {
      const int cj = scan_all ? q : (int)l.cand[ci * DPM2D_MAXCAND + q];
      if (cj == ci) continue;
      over = crossing_odd(l, cj, px, py, T.pbc, L);
    }
    if (over) {
      float dx = l.comx[ci] - px, dy = l.comy[ci] - py;
      if (T.pbc) {
        dx -= L * roundf(dx / L);
        dy -= L * roundf(dy / L);
      }
      const float dist = sqrtf(dx * dx + dy * dy);
      const float xij = dist / (2.0f * a.r0[gc]);
      const float ftmp = T.Kre * (1.0f - xij);
      float ux, uy;
      normalize2(dx, dy, ux, uy);
      fx += 0.5f * ftmp * ux;
      fy += 0.5f * ftmp * uy;
    }
    *n_over = over ? 1 : 0;
  }
  Fx = fx;
  Fy = fy;
}

extern "C" __global__ void __launch_bounds__(1024) k2d_run(const Run2D a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int t = blockIdx.x;
  Tissue2DCtx T;
  T.c0 = a.tissue_cell_begin[t];
  T.nc = a.tissue_cell_begin[t + 1] - T.c0;
  T.v0 = a.cell_vert_begin[T.c0];
  T.nvt = a.cell_vert_begin[T.c0 + T.nc] - T.v0;
  T.L = a.tL[t];
  T.pbc = a.tpbc[t];
  T.Kre = a.tKre[t];
  T.Kat = a.tKat[t];
  if ((a.mask & M2_REPULSION) == 0) T.Kre = 0.0f;
  Lds2D l = carve2d(smem, T.nvt, T.nc);
  const int tid = threadIdx.x, NT = blockDim.x;

  for (int c = tid; c <= T.nc; c += NT) l.vb[c] = a.cell_vert_begin[T.c0 + c] - T.v0;
  for (int k = tid; k < T.nvt; k += NT) {
    l.x[k] = a.x[T.v0 + k];
    l.y[k] = a.y[T.v0 + k];
  }
  __syncthreads();
  for (int c = tid; c < T.nc; c += NT)
    for (int k = l.vb[c]; k < l.vb[c + 1]; k++) l.vcell[k] = (short)c;
  __syncthreads();

  const int nsteps = a.nsteps;
  const int nevals = nsteps > 0 ? nsteps : 1;
  unsigned long long n_over_total = 0;
  float rfx[DPM2D_VPT], rfy[DPM2D_VPT];
  for (int s = 0; s < nevals; s++) {
    cell_pass(a, T, l);
    __syncthreads();
    cand_pass(a, T, l);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < DPM2D_VPT; r++) {
      const int k = tid + r * NT;
      rfx[r] = 0.0f;
      rfy[r] = 0.0f;
      if (k < T.nvt) {
        int no = 0, na = 0;
        force_vertex(a, T, l, k, rfx[r], rfy[r], &no, &na);
        n_over_total += no;
        if (a.p_overlap) a.p_overlap[T.v0 + k] = no;
        if (a.p_attr) a.p_attr[T.v0 + k] = na;
      }
    }
    __syncthreads();
    if (nsteps > 0) {
      const float dt = a.dt;
#pragma unroll
      for (int r = 0; r < DPM2D_VPT; r++) {
        const int k = tid + r * NT;
        if (k < T.nvt) {
          l.x[k] = fmaf(rfx[r], dt, l.x[k]);
          l.y[k] = fmaf(rfy[r], dt, l.y[k]);
        }
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int r = 0; r < DPM2D_VPT; r++) {
    const int k = tid + r * NT;
    if (k < T.nvt) {
      a.x[T.v0 + k] = l.x[k];
      a.y[T.v0 + k] = l.y[k];
      a.fx[T.v0 + k] = rfx[r];
      a.fy[T.v0 + k] = rfy[r];
    }
  }
  if (a.counters) {
    const double c = block_sum((double)n_over_total, l.red);
    if (tid == 0) {
      unsigned long long* slot = a.counters + 4 * (blockIdx.x % DPM_COUNTER_SLOTS);
      if (blockIdx.x == 0) atomicAdd(&slot[0], (unsigned long long)nevals);
      if (c != 0.0) atomicAdd(&slot[2], (unsigned long long)c);
    }
  }
}
