// 2D DPM kernels for gfx950.
//
// A 2D tissue is small enough to live entirely in one workgroup's LDS, so a
// whole CLEulerUpdate call (nsteps steps) — or a whole FIRE minimisation, or
// a whole jamming protocol — is ONE launch with one workgroup per tissue: a
// batch of independent tissues of uneven size runs as one grid.  Steps are
// separated by workgroup barriers only; positions go back to HBM at the end.
//
// Reference semantics followed (Cell2D_kernel.cl):
//   AreaForceUpdates       :13-42   (area summed in double per cell)
//   PerimeterForceUpdates  :92-119
//   BendingForceUpdates    :44-90
//   AttractionForceUpdate  :222-268 (vertex pairs closer than l0 of own cell)
//   RepulsionForceUpdate   :121-220 (crossing parity with its PBC quirk)
//   EulerUpdate            :270-281
// Pair predicates (d < l0, crossing parity) are evaluated with the
// reference's operation order; cells are skipped only where the predicate is
// provably false for every vertex pair (guards below).  Two terms have no
// reference counterpart and make the force law conservative for FIRE: the
// area force with the gradient's y sign, and a vertex soft-disk repulsion.
//
// The real type R is float (the reference arithmetic, with its double
// promotions) or double (the same formulas in double).
#include "dpm_device.h"

using namespace dpm;

#define DPM2D_MAXCAND 48
#define DPM2D_RED 88  // doubles of reduction scratch
#define DPM2D_CSPLIT 8  // threads per cell in the candidate pass (chunks of <= 32 cells)

enum : unsigned {
  M2_AREA = 1u,
  M2_PERIMETER = 2u,
  M2_BENDING = 4u,
  M2_ATTRACTION = 8u,
  M2_REPULSION = 16u,
  M2_AREA_GRAD = 32u,
  M2_VREPULSION = 64u,
};

// Per-tissue geometry and parameters, shared by the Euler and relax kernels.
template <typename R>
struct Geom2D {
  const int* tissue_cell_begin;  // [T+1]
  const int* cell_vert_begin;    // [C+1] (global vertex offsets)
  const float* Ka;
  const float* Kl;
  const float* Kb;
  const float* a0;
  const float* l0;
  const float* r0;
  const R* tL;
  const int* tpbc;
  const float* tKre;
  const float* tKat;
  unsigned mask;
};

// LDS per tissue: x, y [nvt] R | vb [nc+1] int | per cell strain, orient,
// comx, comy, ax0, ax1, ay0, ay1, l0 [nc] R | ncand [nc] int |
// cand [nc*MAXCAND] short | reduction scratch.
__host__ __device__ inline size_t lds2d_bytes(int nvt, int nc, int rs) {
  size_t b = (size_t)rs * 2u * (size_t)nvt;
  b += sizeof(int) * (size_t)(nc + 1);
  b = (b + 7u) & ~(size_t)7u;
  b += (size_t)rs * 9u * (size_t)nc + sizeof(int) * (size_t)nc;
  b += sizeof(short) * (size_t)nc * DPM2D_MAXCAND;
  b = (b + 3u) & ~(size_t)3u;
  b += sizeof(unsigned) * (size_t)nc * DPM2D_CSPLIT;
  b = (b + 7u) & ~(size_t)7u;
  return b + sizeof(double) * DPM2D_RED;
}

template <typename R>
struct Lds2D {
  R *x, *y;
  int* vb;
  R *strain, *orient, *comx, *comy, *ax0, *ax1, *ay0, *ay1;
  R* l0c;  // the cells' l0 (staged from the parameters; the pair guards read it per pair)
  int* ncand;
  short* cand;
  unsigned* cbits;  // [nc * CSPLIT] candidate bits of one chunk of cells
  double* red;
};

template <typename R>
__device__ __forceinline__ Lds2D<R> carve2d(unsigned char* smem, int nvt, int nc) {
  Lds2D<R> l;
  l.x = reinterpret_cast<R*>(smem);
  l.y = l.x + nvt;
  size_t o = sizeof(R) * 2u * (size_t)nvt;
  l.vb = reinterpret_cast<int*>(smem + o);
  o += sizeof(int) * (size_t)(nc + 1);
  o = (o + 7u) & ~(size_t)7u;
  R* cf = reinterpret_cast<R*>(smem + o);
  l.strain = cf;
  l.orient = cf + nc;
  l.comx = cf + 2 * nc;
  l.comy = cf + 3 * nc;
  l.ax0 = cf + 4 * nc;
  l.ax1 = cf + 5 * nc;
  l.ay0 = cf + 6 * nc;
  l.ay1 = cf + 7 * nc;
  l.l0c = cf + 8 * nc;
  o += sizeof(R) * 9u * (size_t)nc;
  l.ncand = reinterpret_cast<int*>(smem + o);
  o += sizeof(int) * (size_t)nc;
  l.cand = reinterpret_cast<short*>(smem + o);
  o += sizeof(short) * (size_t)nc * DPM2D_MAXCAND;
  o = (o + 3u) & ~(size_t)3u;
  l.cbits = reinterpret_cast<unsigned*>(smem + o);
  o += sizeof(unsigned) * (size_t)nc * DPM2D_CSPLIT;
  o = (o + 7u) & ~(size_t)7u;
  l.red = reinterpret_cast<double*>(smem + o);
  return l;
}

template <typename R>
struct Tissue2DCtx {
  int c0, nc, v0, nvt;  // first cell, cells, first vertex, vertices
  R L;
  int pbc;
  R Kre, Kat;
};

template <typename R>
__device__ __forceinline__ Tissue2DCtx<R> tissue_ctx(const Geom2D<R>& g, int t) {
  Tissue2DCtx<R> T;
  T.c0 = g.tissue_cell_begin[t];
  T.nc = g.tissue_cell_begin[t + 1] - T.c0;
  T.v0 = g.cell_vert_begin[T.c0];
  T.nvt = g.cell_vert_begin[T.c0 + T.nc] - T.v0;
  T.L = g.tL[t];
  T.pbc = g.tpbc[t];
  T.Kre = (R)g.tKre[t];
  T.Kat = (R)g.tKat[t];
  return T;
}

// Cell owning tissue-local vertex k: the last c with vb[c] <= k.
__device__ __forceinline__ int cell_of(const int* vb, int nc, int k) {
  int lo = 0, hi = nc - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (vb[mid] <= k) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Per-cell quantities of the current positions: area strain (double area),
// orientation, centre of mass, axis-aligned bounds, of tissue-local cell c.
template <typename R>
__device__ __forceinline__ void cell_one(const Geom2D<R>& g, const Tissue2DCtx<R>& T, const Lds2D<R>& l, int c) {
  const int b = l.vb[c], e = l.vb[c + 1], n = e - b;
  double area = 0.0, sx = 0.0, sy = 0.0;
  R x0 = l.x[b], x1 = l.x[b], y0 = l.y[b], y1 = l.y[b];
  // vertex k's predecessor (the last vertex before the first) carried in
  // registers, the loads of several vertices issued ahead: the same
  // operations in the same order as one vertex after the other
  R xp = l.x[e - 1], yp = l.y[e - 1];
#pragma unroll 4
  for (int j = 0; j < n; j++) {
    const int k = b + j;
    const R xk = l.x[k], yk = l.y[k];
    const double xs = (double)xp + (double)xk;
    const double yd = (double)yp - (double)yk;
    area += xs * yd;
    sx += (double)xk;
    sy += (double)yk;
    x0 = fmin(x0, xk);
    x1 = fmax(x1, xk);
    y0 = fmin(y0, yk);
    y1 = fmax(y1, yk);
    xp = xk;
    yp = yk;
  }
  l.strain[c] = (R)(0.5 * fabs(area) / (double)g.a0[T.c0 + c] - 1.0);
  // the shoelace sum in this order is negative for counter-clockwise cells
  l.orient[c] = area < 0.0 ? R(-1) : R(1);
  l.comx[c] = (R)(sx / n);
  l.comy[c] = (R)(sy / n);
  l.ax0[c] = x0;
  l.ax1[c] = x1;
  l.ay0[c] = y0;
  l.ay1[c] = y1;
}

// cell_one for every cell of the tissue, one thread per cell.
template <typename R>
__device__ void cell_pass(const Geom2D<R>& g, const Tissue2DCtx<R>& T, const Lds2D<R>& l) {
  for (int c = threadIdx.x; c < T.nc; c += blockDim.x) cell_one(g, T, l, c);
}

// Cell-level candidate lists: cell cj is a candidate of ci unless every
// vertex pair is provably outside every pair predicate.
//  - distance predicates (attraction d < l0[ci], vertex repulsion
//    d < (l0[ci] + l0[cj])/2, both under the per-pair minimum image): per
//    axis the minimum-image gap between the two boxes exceeds the cut-off
//    and the boxes are small against L (then no other image comes closer);
//  - repulsion (crossing parity of raw differences, transformed only where a
//    difference exceeds L): the raw boxes are disjoint and no raw
//    difference can exceed L (then the parity of a point outside cj's box
//    is even).
template <typename R>
__device__ bool maybe_pair(const Geom2D<R>& g, const Tissue2DCtx<R>& T, const Lds2D<R>& l, int ci, int cj) {
  const R L = T.L;
  const R tiny = R(1e-5f);
  const R spanx = fmax(l.ax1[ci] - l.ax0[cj], l.ax1[cj] - l.ax0[ci]);
  const R spany = fmax(l.ay1[ci] - l.ay0[cj], l.ay1[cj] - l.ay0[ci]);
  if (g.mask & M2_REPULSION) {
    const bool raw_guard = !T.pbc || (spanx < L * R(0.999f) && spany < L * R(0.999f));
    if (!raw_guard) return true;
    const R m = tiny * (R(1) + fmax(fabs(l.ax1[cj]), fabs(l.ax0[cj])) + fmax(fabs(l.ay1[cj]), fabs(l.ay0[cj])));
    const bool disjoint = l.ax1[ci] < l.ax0[cj] - m || l.ax0[ci] > l.ax1[cj] + m ||
                          l.ay1[ci] < l.ay0[cj] - m || l.ay0[ci] > l.ay1[cj] + m;
    if (!disjoint) return true;
  }
  if (g.mask & (M2_ATTRACTION | M2_VREPULSION)) {
    const R l0i = l.l0c[ci], l0j = l.l0c[cj];
    const R cut = (g.mask & M2_VREPULSION) ? fmax(l0i, l0j) : l0i;
    const R hx = R(0.5f) * ((l.ax1[ci] - l.ax0[ci]) + (l.ax1[cj] - l.ax0[cj]));
    const R hy = R(0.5f) * ((l.ay1[ci] - l.ay0[ci]) + (l.ay1[cj] - l.ay0[cj]));
    R dx = R(0.5f) * (l.ax0[cj] + l.ax1[cj]) - R(0.5f) * (l.ax0[ci] + l.ax1[ci]);
    R dy = R(0.5f) * (l.ay0[cj] + l.ay1[cj]) - R(0.5f) * (l.ay0[ci] + l.ay1[ci]);
    if (T.pbc) {
      if (!(hx + cut < R(0.45f) * L && hy + cut < R(0.45f) * L)) return true;
      dx -= L * round(dx / L);
      dy -= L * round(dy / L);
    }
    const R reach = cut * R(1.0001f) + tiny * (R(1) + L);
    const bool far = fabs(dx) - hx > reach || fabs(dy) - hy > reach;
    if (!far) return true;
  }
  return false;
}

// The candidate list of tissue-local cell ci, ascending, by one thread.
template <typename R>
__device__ __forceinline__ void cand_one(const Geom2D<R>& g, const Tissue2DCtx<R>& T, const Lds2D<R>& l, int ci) {
  int n = 0;
  for (int cj = 0; cj < T.nc; cj++) {
    if (cj == ci) continue;
    if (maybe_pair(g, T, l, ci, cj)) {
      if (n < DPM2D_MAXCAND) l.cand[ci * DPM2D_MAXCAND + n] = (short)cj;
      n++;
    }
  }
  l.ncand[ci] = n;  // > MAXCAND: the vertex loops scan every cell
}

template <typename R>
__device__ void cand_pass(const Geom2D<R>& g, const Tissue2DCtx<R>& T, const Lds2D<R>& l) {
  const int nc = T.nc, NT = blockDim.x;
  const int S = max(1, min(DPM2D_CSPLIT, NT / max(1, nc)));  // threads per cell
  const int chunk = (nc + S - 1) / S;
  if (chunk <= 32) {  // block-uniform: each thread tests one chunk of <= 32 cells into a bit mask
    for (int w = threadIdx.x; w < nc * S; w += NT) {
      const int ci = w % nc, sp = w / nc;
      unsigned bits = 0u;
      const int c1 = min(nc, (sp + 1) * chunk);
      for (int cj = sp * chunk; cj < c1; cj++)
        if (cj != ci && maybe_pair(g, T, l, ci, cj)) bits |= 1u << (cj - sp * chunk);
      l.cbits[ci * DPM2D_CSPLIT + sp] = bits;
    }
    __syncthreads();
    for (int ci = threadIdx.x; ci < nc; ci += NT) {  // the list in ascending cell order
      int n = 0;
      for (int sp = 0; sp < S; sp++) {
        unsigned bits = l.cbits[ci * DPM2D_CSPLIT + sp];
        while (bits) {
          const int cj = sp * chunk + __ffs((int)bits) - 1;
          bits &= bits - 1u;
          if (n < DPM2D_MAXCAND) l.cand[ci * DPM2D_MAXCAND + n] = (short)cj;
          n++;
        }
      }
      l.ncand[ci] = n;  // > MAXCAND: the vertex loops scan every cell
    }
    return;
  }
  for (int ci = threadIdx.x; ci < nc; ci += NT) cand_one(g, T, l, ci);
}

// True when no vertex of cell cj can lie within `cut` of point p under the
// per-pair minimum image: per axis the minimum-image gap between p and cj's
// axis-aligned box exceeds the cut-off, and the box is small against L (then
// no other image comes closer) — maybe_pair's argument with a point for the
// first box.  Skipping such a cell leaves every sum unchanged.
template <typename R>
__device__ __forceinline__ bool vertex_far_from_box(const Lds2D<R>& l, const Tissue2DCtx<R>& T, int cj, R px, R py,
                                                    R cut) {
  const R L = T.L;
  const R hx = R(0.5f) * (l.ax1[cj] - l.ax0[cj]), hy = R(0.5f) * (l.ay1[cj] - l.ay0[cj]);
  R dx = R(0.5f) * (l.ax0[cj] + l.ax1[cj]) - px, dy = R(0.5f) * (l.ay0[cj] + l.ay1[cj]) - py;
  if (T.pbc) {
    if (!(hx + cut < R(0.45f) * L && hy + cut < R(0.45f) * L)) return false;
    dx -= L * round(dx / L);
    dy -= L * round(dy / L);
  }
  const R reach = cut * R(1.0001f) + R(1e-5f) * (R(1) + L);
  return fabs(dx) - hx > reach || fabs(dy) - hy > reach;
}

// floor(x / L) and round(x / L) as the reference evaluates them (a
// correctly rounded quotient), from q = x (1/L) where that is certain: q is
// within |q| 2^-22 of the rounded quotient, so where q is more than
// 1e-6 (1 + |q|) away from every integer (for round, every half-integer)
// both give the same result; closer, the division is taken.
template <typename R>
__device__ __forceinline__ R floor_quot(R x, R L, R invL) {
  if constexpr (sizeof(R) == 4) {
    const R q = x * invL, f = floor(q), m = R(1e-6f) * (R(1) + fabs(q));
    if (q - f > m && (f + R(1)) - q > m) return f;
  }
  return floor(x / L);
}
template <typename R>
__device__ __forceinline__ R round_quot(R x, R L, R invL) {
  if constexpr (sizeof(R) == 4) {
    const R q = x * invL, m = R(1e-6f) * (R(1) + fabs(q));
    const R h = fabs(q) - floor(fabs(q));  // distance above the integer below |q|
    if (fabs(h - R(0.5f)) > m) return round(q);
  }
  return round(x / L);
}

// Crossing parity of point p against cell cj, Cell2D_kernel.cl:171-196.
// The crossing abscissa tt = (djx - dix)(0 - diy)/(djy - diy) + dix only
// decides 0 < tt: for float its sign is taken from the product with the
// hardware reciprocal (relative error below 2^-21) where |tt| exceeds
// 1e-6 of its terms, else from the correctly rounded division.  (The
// rounded sum of two floats is 0 only when the exact sum is, so the sign of
// tt is the sign of q + dix for the rounded quotient q.)
template <typename R>
__device__ __forceinline__ bool crossing_odd(const Lds2D<R>& l, int cj, R px, R py, int pbc, R L,
                                             int first = 0, int stride = 1) {
  // the edges (i, i - 1) for i = first, first + stride, ... (all of them by
  // default; a share of them when several lanes split one polygon)
  const int b = l.vb[cj], e = l.vb[cj + 1], n = e - b;
  const R invL = R(1) / L;
  bool odd = false;
  for (int i = first; i < n; i += stride) {
    const int j = i == 0 ? n - 1 : i - 1;
    R dix = px - l.x[b + i], diy = py - l.y[b + i];
    R djx = px - l.x[b + j], djy = py - l.y[b + j];
    if (pbc) {
      if (fabs(dix) > L || fabs(djx) > L) {
        dix -= L * floor_quot(dix, L, invL);
        djx -= L * floor_quot(djx, L, invL);
      }
      if (fabs(diy) > L || fabs(djy) > L) {
        diy -= L * round_quot(diy, L, invL);
        djy -= L * round_quot(djy, L, invL);
      }
    }
    if ((diy > R(0)) != (djy > R(0))) {
      const R a = (djx - dix) * (R(0) - diy), c = djy - diy;
      bool pos;
      if constexpr (sizeof(R) == 4) {
        const R qa = a * __builtin_amdgcn_rcpf(c), ta = qa + dix;
        if (fabs(ta) > R(1e-6f) * (fabs(qa) + fabs(dix)))
          pos = ta > R(0);
        else
          pos = R(0) < a / c + dix;
      } else {
        pos = R(0) < a / c + dix;
      }
      if (pos) odd = !odd;
    }
  }
  return odd;
}

// A point outside cell cj's raw bounding box by more than 1e-5 (1 + L),
// whose raw differences to every vertex of cj stay within L (so crossing_odd
// applies no image), has even crossing parity: above or below the box no
// edge straddles its row; left of the box every crossing lies to its right
// and none counts; right of the box every one counts, and a closed ring has
// an even number of edges straddling any row.
template <typename R>
__device__ __forceinline__ bool outside_box_even(const Lds2D<R>& l, int cj, R px, R py, int pbc, R L) {
  const R x0 = l.ax0[cj], x1 = l.ax1[cj], y0 = l.ay0[cj], y1 = l.ay1[cj];
  if (pbc && !(fabs(px - x0) <= L && fabs(px - x1) <= L && fabs(py - y0) <= L && fabs(py - y1) <= L))
    return false;
  const R m = R(1e-5f) * (R(1) + L);
  return px < x0 - m || px > x1 + m || py < y0 - m || py > y1 + m;
}

// Crossing parity of every vertex against every candidate cell: bit k of
// `bits` (cleared beforehand) is set when vertex k lies inside some
// candidate by the odd-crossing rule (Cell2D_kernel.cl:165-202).  That is
// the OR force_vertex forms over the same candidates, so the two agree bit
// for bit.  First one (vertex, candidate) pair per thread and iteration
// takes the cheap tests (a candidate slot in range, not the own cell, not
// outside the candidate's box); the pairs that pass are appended to `list`
// (wave-aggregated LDS atomics on *count, zeroed beforehand), and after a
// workgroup barrier the crossing loops run on the list over full waves —
// left in place, the few passing pairs of each wave would hold all its
// lanes through their loops.  Pairs beyond `cap` entries (cap 0: no list)
// are tested where they are found.  The bits are an OR, whatever the order.
// Slots q run below qn, the largest candidate count of the tissue's cells
// (nc where a list overflowed), and (k, q) = (w mod nvt, w / nvt) is
// stepped without divisions.
template <typename R>
__device__ void over_pass(const Tissue2DCtx<R>& T, const Lds2D<R>& l, unsigned* bits, int* list, int cap,
                          int* count, int qn) {
  const int nvt = T.nvt, nc = T.nc, lane = threadIdx.x & (DPM_WAVE - 1), NT = blockDim.x;
  const int dq = NT / nvt, dk = NT - dq * nvt;
  int k = (int)threadIdx.x % nvt, q = (int)threadIdx.x / nvt;
  for (int w = threadIdx.x; w < nvt * qn; w += NT, k += dk, q += dq) {
    if (k >= nvt) {
      k -= nvt;
      q++;
    }
    const int ci = cell_of(l.vb, nc, k);
    const int ncd = l.ncand[ci];
    const bool scan_all = ncd > DPM2D_MAXCAND;
    bool pass = false;
    int cj = 0;
    R px = R(0), py = R(0);
    if (q < (scan_all ? nc : ncd)) {
      cj = scan_all ? q : (int)l.cand[ci * DPM2D_MAXCAND + q];
      px = l.x[k];
      py = l.y[k];
      pass = cj != ci && !outside_box_even(l, cj, px, py, T.pbc, T.L);
    }
    if (cap == 0) {  // block-uniform: no list (count and list are not allocated)
      if (pass && crossing_odd(l, cj, px, py, T.pbc, T.L)) atomicOr(&bits[k >> 5], 1u << (k & 31));
      continue;
    }
    const unsigned long long m = __ballot(pass);
    if (m == 0ull) continue;
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(count, (int)__popcll(m));
    base = __shfl(base, leader);
    if (pass) {
      const int slot = base + (int)__popcll(m & ((1ull << lane) - 1ull));
      if (slot < cap)
        list[slot] = k | (cj << 16);
      else if (crossing_odd(l, cj, px, py, T.pbc, T.L))
        atomicOr(&bits[k >> 5], 1u << (k & 31));
    }
  }
  if (cap == 0) return;
  __syncthreads();
  // eight lanes per listed pair (aligned groups, active together), each
  // testing every eighth edge; the crossing parity is the XOR over the group
  const int n = min(*count, cap);
  for (int i = threadIdx.x; i < 8 * n; i += blockDim.x) {
    const int e = list[i >> 3], g = i & 7, k = e & 0xffff, cj = e >> 16;
    int odd = crossing_odd(l, cj, l.x[k], l.y[k], T.pbc, T.L, g, 8) ? 1 : 0;
    odd ^= __shfl_xor(odd, 1);
    odd ^= __shfl_xor(odd, 2);
    odd ^= __shfl_xor(odd, 4);
    if (g == 0 && odd) atomicOr(&bits[k >> 5], 1u << (k & 31));
  }
}

// All force terms on vertex k (tissue-local index).  With `energy`, also
// accumulates this vertex's share of the conservative energy (perimeter edge
// k -> k+1, bending at k, the cell's area energy at its first vertex) and
// with `e_int` half of each vertex-repulsion pair energy.  With `over_bits`
// the crossing parity comes from over_pass instead of the loop here, with
// over_state (0 or 1) from k2d_wide_over.
template <typename R>
__device__ __forceinline__ void force_vertex(const Geom2D<R>& g, const Tissue2DCtx<R>& T, const Lds2D<R>& l,
                                          int k, R& Fx, R& Fy, int* n_over, int* n_attr, double* energy,
                                          double* e_int, const unsigned* over_bits = nullptr,
                                          unsigned* attr_key = nullptr, int over_state = -1) {
  const int ci = cell_of(l.vb, T.nc, k);
  const int gc = T.c0 + ci;
  const int b = l.vb[ci], e = l.vb[ci + 1], n = e - b, vi = k - b;
  const int im1 = (vi == 0) ? e - 1 : k - 1;
  const int ip1 = (vi == n - 1) ? b : k + 1;
  const R px = l.x[k], py = l.y[k];
  const R L = T.L;
  R fx = R(0), fy = R(0);
  const unsigned mask = g.mask;
  if (mask & M2_AREA) {
    const R Ka = (R)g.Ka[gc], a0 = (R)g.a0[gc];
    const double pre = (double)(Ka / sqrt(a0)) * 0.5 * (double)l.strain[ci];
    fx = (R)((double)fx + pre * (double)(l.y[im1] - l.y[ip1]));
    fy = (R)((double)fy + pre * (double)(l.x[im1] - l.x[ip1]));
  }
  if (mask & M2_AREA_GRAD) {
    // -dE/dx of E = Ka sqrt(a0)/2 s^2, s = |A|/a0 - 1, either orientation
    const R Ka = (R)g.Ka[gc], a0 = (R)g.a0[gc];
    const R c = Ka / sqrt(a0) * l.strain[ci] * l.orient[ci] * R(0.5f);
    fx -= c * (l.y[im1] - l.y[ip1]);
    fy -= c * (l.x[ip1] - l.x[im1]);
  }
  if (energy && vi == 0 && (mask & (M2_AREA | M2_AREA_GRAD))) {
    const double s = (double)l.strain[ci];
    *energy += 0.5 * (double)g.Ka[gc] * sqrt((double)g.a0[gc]) * s * s;
  }
  if (mask & M2_PERIMETER) {
    const R l0 = (R)g.l0[gc];
    const R lvx = l.x[ip1] - px, lvy = l.y[ip1] - py;
    const R lvmx = px - l.x[im1], lvmy = py - l.y[im1];
    const R len = sqrt(lvx * lvx + lvy * lvy);
    const R lenm = sqrt(lvmx * lvmx + lvmy * lvmy);
    const R dli = (R)((double)(len / l0) - 1.0);
    const R dlim1 = (R)((double)(lenm / l0) - 1.0);
    const R pref = (R)g.Kl[gc] * sqrt((R)g.a0[gc] / l0);
    fx += pref * (dli * (lvx / len) - dlim1 * (lvmx / lenm));
    fy += pref * (dli * (lvy / len) - dlim1 * (lvmy / lenm));
    if (energy) {
      const double dl0 = (double)g.l0[gc];
      const double d = sqrt((double)lvx * lvx + (double)lvy * lvy) / dl0 - 1.0;
      *energy += 0.5 * (double)g.Kl[gc] * sqrt((double)g.a0[gc] / dl0) * dl0 * d * d;
    }
  }
  if (mask & M2_BENDING) {
    const int ip2 = (vi == n - 1) ? b + 1 : ((vi == n - 2) ? b : k + 2);
    const int im2 = (vi == 0) ? e - 2 : ((vi == 1) ? e - 1 : k - 2);
    const R lvx = l.x[ip1] - px, lvy = l.y[ip1] - py;
    const R lvxm = px - l.x[im1], lvym = py - l.y[im1];
    const R six = lvx - lvxm, siy = lvy - lvym;
    const R sixp = (l.x[ip2] - l.x[ip1]) - lvx, siyp = (l.y[ip2] - l.y[ip1]) - lvy;
    const R sixm = lvxm - (l.x[im1] - l.x[im2]), siym = lvym - (l.y[im1] - l.y[im2]);
    const double Kb = (double)g.Kb[gc];
    fx = (R)((double)fx + Kb * (2.0 * (double)six - (double)sixm - (double)sixp));
    fy = (R)((double)fy + Kb * (2.0 * (double)siy - (double)siym - (double)siyp));
    if (energy) {
      const double sx = ((double)l.x[ip1] - (double)px) - ((double)px - (double)l.x[im1]);
      const double sy = ((double)l.y[ip1] - (double)py) - ((double)py - (double)l.y[im1]);
      *energy += 0.5 * Kb * (sx * sx + sy * sy);
    }
  }
  const int nc = l.ncand[ci];
  const bool scan_all = nc > DPM2D_MAXCAND;
  const int ncl = scan_all ? T.nc : nc;
  const R half_l = R(0.49f) * L;
  // With Kat = 0 the attraction adds exact zeros: it is evaluated only where
  // its partners are asked for (n_attr, attr_key).
  if ((mask & M2_ATTRACTION) && (T.Kat != R(0) || n_attr || attr_key)) {
    const R l0 = (R)g.l0[gc];
    const R kat = T.Kat / (R)n;
    const R lim = l0 * l0 * R(1.00001f);
    const double kk = (double)kat / (double)l0;  // ftmp/d = kat/l0 of an accepted pair
    int cnt = 0;
    for (int q = 0; q < ncl; q++) {
      const int cj = scan_all ? q : (int)l.cand[ci * DPM2D_MAXCAND + q];
      if (cj == ci || vertex_far_from_box(l, T, cj, px, py, l0)) continue;
      for (int j = l.vb[cj]; j < l.vb[cj + 1]; j++) {
        R rx = l.x[j] - px, ry = l.y[j] - py;
        R nx = R(0), ny = R(0);
        // minimum image; a component below 0.49 L rounds to image 0, which
        // leaves it unchanged, so the division is skipped there
        if (T.pbc && !(fabs(rx) < half_l && fabs(ry) < half_l)) {
          nx = round(rx / L);
          ny = round(ry / L);
          rx -= L * nx;
          ry -= L * ny;
        }
        const R d2 = rx * rx + ry * ry;
        if (!(d2 < lim)) continue;  // sqrt(d2) < l0 needs d2 < l0^2 (1 + 1e-5)
        const R dist = sqrt(d2);
        if (dist < l0) {
          // the reference's decision; the force (kat d/l0 along r) in
          // double on the same image, as a short image vector keeps few
          // digits in float
          const double ddx = ((double)l.x[j] - (double)px) - (double)L * (double)nx;
          const double ddy = ((double)l.y[j] - (double)py) - (double)L * (double)ny;
          fx += (R)(kk * ddx);
          fy += (R)(kk * ddy);
          cnt++;
          if (attr_key) atomicAdd(attr_key, vertex_key(l.x[j], l.y[j]));
        }
      }
    }
    if (n_attr) *n_attr = cnt;
  }
  if (mask & M2_VREPULSION) {
    // U = Kre/2 (1 - d/sigma)^2 for d < sigma = (l0_i + l0_j)/2
    const R l0i = (R)g.l0[gc];
    for (int q = 0; q < ncl; q++) {
      const int cj = scan_all ? q : (int)l.cand[ci * DPM2D_MAXCAND + q];
      if (cj == ci) continue;
      const R sig = R(0.5f) * (l0i + (R)g.l0[T.c0 + cj]);
      if (vertex_far_from_box(l, T, cj, px, py, sig)) continue;
      for (int j = l.vb[cj]; j < l.vb[cj + 1]; j++) {
        R rx = l.x[j] - px, ry = l.y[j] - py;
        if (T.pbc && !(fabs(rx) < half_l && fabs(ry) < half_l)) {
          rx -= L * round(rx / L);
          ry -= L * round(ry / L);
        }
        const R d2 = rx * rx + ry * ry;
        if (!(d2 < sig * sig) || d2 == R(0)) continue;
        const R dist = sqrt(d2);
        const R ov = R(1) - dist / sig;
        const R ft = T.Kre / sig * ov / dist;
        fx -= ft * rx;
        fy -= ft * ry;
        if (e_int) *e_int += 0.25 * (double)T.Kre * (double)ov * (double)ov;  // half of the pair
      }
    }
  }
  if (mask & M2_REPULSION) {
    // the crossing parity given (over_state 0 or 1, or bit k of over_bits),
    // else found by the loop
    const bool given = over_state >= 0 || over_bits;
    bool over = over_state >= 0 ? over_state != 0 : (over_bits && ((over_bits[k >> 5] >> (k & 31)) & 1u));
    for (int q = 0; q < ncl && !over && !given; q++) {
      const int cj = scan_all ? q : (int)l.cand[ci * DPM2D_MAXCAND + q];
      if (cj == ci || outside_box_even(l, cj, px, py, T.pbc, L)) continue;
      over = crossing_odd(l, cj, px, py, T.pbc, L);
    }
    if (over) {
      R dx = l.comx[ci] - px, dy = l.comy[ci] - py;
      if (T.pbc) {
        dx -= L * round(dx / L);
        dy -= L * round(dy / L);
      }
      const R dist = sqrt(dx * dx + dy * dy);
      const R xij = dist / (R(2) * (R)g.r0[gc]);
      const R ftmp = T.Kre * (R(1) - xij);
      const V3<R> u = vnormalize(V3<R>{dx, dy, R(0)});
      fx += R(0.5f) * ftmp * u.x;
      fy += R(0.5f) * ftmp * u.y;
    }
    *n_over = over ? 1 : 0;
  }
  Fx = fx;
  Fy = fy;
}

template <typename R>
__device__ __forceinline__ void stage_tissue(const Geom2D<R>& g, const Tissue2DCtx<R>& T, const Lds2D<R>& l,
                                             const R* x, const R* y) {
  for (int c = threadIdx.x; c <= T.nc; c += blockDim.x) {
    l.vb[c] = g.cell_vert_begin[T.c0 + c] - T.v0;
    if (c < T.nc) l.l0c[c] = (R)g.l0[T.c0 + c];
  }
  for (int k = threadIdx.x; k < T.nvt; k += blockDim.x) {
    l.x[k] = x[T.v0 + k];
    l.y[k] = y[T.v0 + k];
  }
  __syncthreads();
}

// ------------------------------------------------------------ Euler (float)

struct Run2D {
  Geom2D<float> g;
  float* x;  // [nverts] packed cell after cell, all tissues
  float* y;
  float* fx;  // forces of the last evaluation (written at the end)
  float* fy;
  int nsteps;  // Euler steps (0 = evaluate forces only)
  float dt;
  unsigned long long* counters;  // [4 * DPM_COUNTER_SLOTS]
  int* p_overlap;                // diagnostics, per vertex (nullptr normally)
  int* p_attr;
  unsigned* p_attr_key;  // attraction partner-set key sums (vertex_key)
  int over_cap;          // entries of the crossing-parity pair list after the LDS arrays (0: none)
};

// The crossing-parity pair list of k2d_run (a count, the largest candidate
// count, then the entries; the entries pack vertex | cell << 16): its LDS
// after a tissue's Lds2D arrays.
#define DPM2D_OVER_CAP 2048
__host__ __device__ inline size_t lds2d_over_bytes() { return sizeof(int) * (size_t)(DPM2D_OVER_CAP + 2); }

// Forces stay in registers: up to DPM2D_VPT vertices per thread.
#define DPM2D_VPT 16

extern "C" __global__ void __launch_bounds__(1024) k2d_run(const Run2D a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const Geom2D<float>& g = a.g;
  const Tissue2DCtx<float> T = tissue_ctx(g, blockIdx.x);
  const Lds2D<float> l = carve2d<float>(smem, T.nvt, T.nc);
  const int tid = threadIdx.x, NT = blockDim.x;
  stage_tissue(g, T, l, a.x, a.y);

  const int nsteps = a.nsteps;
  const int nevals = nsteps > 0 ? nsteps : 1;
  unsigned long long n_over_total = 0;
  // Crossing parity in its own pass when the (vertex, candidate) pairs are at
  // most 16 per thread and one bit per vertex fits the candidate-bit words
  // (nc * DPM2D_CSPLIT of them, unused between the candidate and the next
  // step's cell pass).
  const bool split_over = (g.mask & M2_REPULSION) && T.nvt <= 32 * DPM2D_CSPLIT * T.nc &&
                          T.nvt * T.nc <= 16 * NT;
  unsigned* obits = l.cbits;
  int* ocount = reinterpret_cast<int*>(smem + lds2d_bytes(T.nvt, T.nc, (int)sizeof(float)));
  int* oqmax = ocount + 1;
  int* olist = ocount + 2;
  if (tid == 0 && a.over_cap > 0) *oqmax = 0;  // (the first step's barriers precede its use)
  float rfx[DPM2D_VPT], rfy[DPM2D_VPT];
  for (int s = 0; s < nevals; s++) {
    cell_pass(g, T, l);
    __syncthreads();
    cand_pass(g, T, l);
    __syncthreads();
    if (split_over) {  // block-uniform; the candidate bits are dead here
      for (int i = tid; i < (T.nvt + 31) / 32; i += NT) obits[i] = 0u;
      if (a.over_cap > 0) {
        if (tid == 0) *ocount = 0;
        for (int c = tid; c < T.nc; c += NT) {
          const int nd = l.ncand[c];
          atomicMax(oqmax, nd > DPM2D_MAXCAND ? T.nc : nd);
        }
      }
      __syncthreads();
      over_pass(T, l, obits, olist, a.over_cap, ocount, a.over_cap > 0 ? *oqmax : T.nc);
      __syncthreads();
      if (tid == 0 && a.over_cap > 0) *oqmax = 0;  // every thread read it before the barrier above
    }
#pragma unroll
    for (int r = 0; r < DPM2D_VPT; r++) {
      const int k = tid + r * NT;
      rfx[r] = 0.0f;
      rfy[r] = 0.0f;
      if (k < T.nvt) {
        int no = 0, na = 0;
        force_vertex(g, T, l, k, rfx[r], rfy[r], &no, a.p_attr ? &na : nullptr, (double*)nullptr, (double*)nullptr,
                     split_over ? obits : nullptr, a.p_attr_key ? a.p_attr_key + T.v0 + k : nullptr);
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

// ------------------------------------------------- relaxation (float/double)

enum : int {
  RELAX_EULER = 0,   // nsteps overdamped Euler steps
  RELAX_FIRE = 1,    // one FIRE minimisation
  RELAX_JAM = 2,     // jamming: compression stages, each a FIRE minimisation
  RELAX_ENERGY = 3,  // forces and energy at the current positions
};

// FIRE parameters (Bitzek et al., PRL 97, 170201 (2006)).
struct FireParams {
  double dt0, dt_max, ftol, f_inc, f_dec, alpha0, f_alpha;
  int max_iter, n_min;
};

// Jamming parameters: packing fraction phi = sum a0 / L^2.
struct JamParams {
  double phi0;     // first packing fraction (must be unjammed)
  double dphi;     // growth step until the first jammed state
  double phi_tol;  // bisection stops when phi_hi - phi_lo < phi_tol
  double u_jam;    // jammed iff vertex-repulsion energy per cell > u_jam
  int max_stages;
};

// Per-tissue results (doubles), RS_COUNT per tissue.
enum : int {
  RS_ITER = 0,     // Euler steps or FIRE iterations (all stages)
  RS_CONVERGED,    // the final state has max |F_v| < ftol
  RS_FMAX,         // max |F_v| at the end
  RS_ENERGY,       // conservative energy at the end
  RS_EINT,         // its vertex-repulsion part
  RS_PHI,          // packing fraction at the end
  RS_PHI_LO,       // jamming: largest unjammed phi
  RS_PHI_HI,       // jamming: smallest jammed phi (0 = none found)
  RS_STAGES,       // jamming stages run
  RS_L,            // box length at the end
  RS_UNCONVERGED,  // minimisations that stopped at max_iter
  RS_STATUS,       // jamming: 0 bracket within phi_tol, 1 phi0 was jammed,
                   // 2 in progress, 3 stopped at max_stages
  RS_PHI_TARGET,   // jamming: packing fraction of the next minimisation
  RS_COUNT
};

template <typename R>
struct Relax2D {
  Geom2D<R> g;
  R* x;  // [nverts] positions (updated)
  R* y;
  R* vx;  // [nverts] FIRE velocities
  R* vy;
  R* fx;  // [nverts] forces at the final positions
  R* fy;
  R* sx;  // [nverts] saved unjammed positions (jamming)
  R* sy;
  R* tLout;              // [T] box length at the end (jamming changes it)
  double* res;           // [T * RS_COUNT]
  const double* sum_a0;  // [T]
  int mode;
  int nsteps;
  R dt;
  FireParams fire;
  JamParams jam;
  int jam_start;   // 1: begin a jamming protocol; 0: continue the one in res
  int jam_stages;  // minimisations this launch may run
};

// Reduction of five doubles (four sums, one max) over the workgroup; every
// thread gets the results.  scratch: DPM2D_RED doubles.
__device__ __forceinline__ void block_reduce5(double& a, double& b, double& c, double& d, double& m,
                                              double* scratch) {
  const int lane = threadIdx.x & (DPM_WAVE - 1), wave = threadIdx.x / DPM_WAVE;
  const int nw = (blockDim.x + DPM_WAVE - 1) / DPM_WAVE;
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, o, DPM_WAVE);
    b += __shfl_down(b, o, DPM_WAVE);
    c += __shfl_down(c, o, DPM_WAVE);
    d += __shfl_down(d, o, DPM_WAVE);
    m = fmax(m, __shfl_down(m, o, DPM_WAVE));
  }
  if (lane == 0) {
    scratch[wave] = a;
    scratch[16 + wave] = b;
    scratch[32 + wave] = c;
    scratch[48 + wave] = d;
    scratch[64 + wave] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sa = 0.0, sb = 0.0, sc = 0.0, sd = 0.0, sm = 0.0;
    for (int w = 0; w < nw; w++) {
      sa += scratch[w];
      sb += scratch[16 + w];
      sc += scratch[32 + w];
      sd += scratch[48 + w];
      sm = fmax(sm, scratch[64 + w]);
    }
    scratch[80] = sa;
    scratch[81] = sb;
    scratch[82] = sc;
    scratch[83] = sd;
    scratch[84] = sm;
  }
  __syncthreads();
  a = scratch[80];
  b = scratch[81];
  c = scratch[82];
  d = scratch[83];
  m = scratch[84];
  __syncthreads();
}

template <typename R>
__device__ __forceinline__ void prepare_forces(const Geom2D<R>& g, const Tissue2DCtx<R>& T, const Lds2D<R>& l) {
  cell_pass(g, T, l);
  __syncthreads();
  cand_pass(g, T, l);
  __syncthreads();
}

// One FIRE minimisation from the current positions, velocities from 0, in
// the order of the FP64 restatement: F(x); stop if max |F_v| < ftol, or —
// with u_clear >= 0 — if the vertex-repulsion energy per cell is at most
// u_clear (`cleared`: no contact network left); power and norms in double;
// velocity mixing or restart; semi-implicit Euler.  Returns the iterations
// taken.
template <typename R>
__device__ int fire_min(const Relax2D<R>& a, const Tissue2DCtx<R>& T, const Lds2D<R>& l, double u_clear,
                        bool& converged, bool& cleared) {
  const int NT = blockDim.x, tid = threadIdx.x;
  const FireParams& p = a.fire;
  for (int k = tid; k < T.nvt; k += NT) {
    a.vx[T.v0 + k] = R(0);
    a.vy[T.v0 + k] = R(0);
  }
  double dt = p.dt0, alpha = p.alpha0;
  int npos = 0, it = 0;
  converged = false;
  cleared = false;
  for (it = 0; it < p.max_iter; it++) {
    prepare_forces(a.g, T, l);
    double P = 0.0, F2 = 0.0, v2 = 0.0, ui = 0.0, fm = 0.0;
    for (int k = tid; k < T.nvt; k += NT) {
      R fx, fy;
      int no = 0;
      force_vertex(a.g, T, l, k, fx, fy, &no, (int*)nullptr, (double*)nullptr, u_clear >= 0.0 ? &ui : (double*)nullptr);
      a.fx[T.v0 + k] = fx;
      a.fy[T.v0 + k] = fy;
      const double vx = (double)a.vx[T.v0 + k], vy = (double)a.vy[T.v0 + k];
      P += (double)fx * vx + (double)fy * vy;
      F2 += (double)fx * fx + (double)fy * fy;
      v2 += vx * vx + vy * vy;
      fm = fmax(fm, (double)fx * fx + (double)fy * fy);  // squared: one square root after the max
    }
    block_reduce5(P, F2, v2, ui, fm, l.red);
    fm = sqrt(fm);  // correctly rounded and monotone: the max of the roots
    if (u_clear >= 0.0 && ui / T.nc <= u_clear) {
      cleared = true;
      break;
    }
    if (fm < p.ftol) {
      converged = true;
      break;
    }
    double mix_v = 1.0, mix_f = 0.0;
    bool zero_v = false;
    if (P > 0.0) {
      mix_v = 1.0 - alpha;
      mix_f = alpha * sqrt(v2 / F2);
      if (npos > p.n_min) {
        dt = fmin(dt * p.f_inc, p.dt_max);
        alpha *= p.f_alpha;
      }
      npos++;
    } else if (P < 0.0) {
      zero_v = true;
      dt *= p.f_dec;
      alpha = p.alpha0;
      npos = 0;
    }
    for (int k = tid; k < T.nvt; k += NT) {
      const size_t gk = (size_t)T.v0 + k;
      const double fx = (double)a.fx[gk], fy = (double)a.fy[gk];
      R vx = zero_v ? R(0) : (R)(mix_v * (double)a.vx[gk] + mix_f * fx);
      R vy = zero_v ? R(0) : (R)(mix_v * (double)a.vy[gk] + mix_f * fy);
      vx = (R)((double)vx + fx * dt);
      vy = (R)((double)vy + fy * dt);
      a.vx[gk] = vx;
      a.vy[gk] = vy;
      l.x[k] = (R)((double)l.x[k] + (double)vx * dt);
      l.y[k] = (R)((double)l.y[k] + (double)vy * dt);
    }
    __syncthreads();
  }
  return it;
}

// Forces (to a.fx/fy), energy and its vertex-repulsion part at the current
// positions, and max |F_v|.
template <typename R>
__device__ void energy_pass(const Relax2D<R>& a, const Tissue2DCtx<R>& T, const Lds2D<R>& l, double& U,
                            double& Uint, double& fmax_out) {
  prepare_forces(a.g, T, l);
  double u = 0.0, ui = 0.0, unused = 0.0, fm = 0.0;
  for (int k = threadIdx.x; k < T.nvt; k += blockDim.x) {
    R fx, fy;
    int no = 0;
    force_vertex(a.g, T, l, k, fx, fy, &no, (int*)nullptr, &u, &ui);
    a.fx[T.v0 + k] = fx;
    a.fy[T.v0 + k] = fy;
    fm = fmax(fm, (double)fx * fx + (double)fy * fy);
  }
  double unused2 = 0.0;
  block_reduce5(u, ui, unused, unused2, fm, l.red);
  fm = sqrt(fm);
  U = u + ui;
  Uint = ui;
  fmax_out = fm;
}

// Moves every cell rigidly so that its centre of mass scales with the box,
// L -> Lnew: x_v += (Lnew/L - 1) com_c.
template <typename R>
__device__ void rescale_box(const Geom2D<R>& g, Tissue2DCtx<R>& T, const Lds2D<R>& l, R Lnew) {
  cell_pass(g, T, l);
  __syncthreads();
  const R s = Lnew / T.L - R(1);
  for (int k = threadIdx.x; k < T.nvt; k += blockDim.x) {
    const int c = cell_of(l.vb, T.nc, k);
    l.x[k] += s * l.comx[c];
    l.y[k] += s * l.comy[c];
  }
  T.L = Lnew;
  __syncthreads();
}

template <typename R>
__device__ void copy_positions(const Tissue2DCtx<R>& T, const R* fx, const R* fy, R* tx, R* ty) {
  for (int k = threadIdx.x; k < T.nvt; k += blockDim.x) {
    tx[k] = fx[k];
    ty[k] = fy[k];
  }
  __syncthreads();
}

template <typename R>
__device__ void k2d_relax_body(const Relax2D<R>& a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const Geom2D<R>& g = a.g;
  Tissue2DCtx<R> T = tissue_ctx(g, blockIdx.x);
  const Lds2D<R> l = carve2d<R>(smem, T.nvt, T.nc);
  const int tid = threadIdx.x, NT = blockDim.x;
  stage_tissue(g, T, l, a.x, a.y);
  double res[RS_COUNT];
  for (int i = 0; i < RS_COUNT; i++) res[i] = 0.0;
  const double A0 = a.sum_a0[blockIdx.x];

  if (a.mode == RELAX_EULER) {
    for (int s = 0; s < a.nsteps; s++) {
      prepare_forces(g, T, l);
      for (int k = tid; k < T.nvt; k += NT) {
        R fx, fy;
        int no = 0;
        force_vertex(g, T, l, k, fx, fy, &no, (int*)nullptr, (double*)nullptr, (double*)nullptr);
        a.fx[T.v0 + k] = fx;
        a.fy[T.v0 + k] = fy;
      }
      __syncthreads();
      for (int k = tid; k < T.nvt; k += NT) {
        l.x[k] = fma(a.fx[T.v0 + k], a.dt, l.x[k]);
        l.y[k] = fma(a.fy[T.v0 + k], a.dt, l.y[k]);
      }
      __syncthreads();
    }
    res[RS_ITER] = a.nsteps;
  } else if (a.mode == RELAX_FIRE) {
    bool conv, cleared;
    res[RS_ITER] = fire_min(a, T, l, -1.0, conv, cleared);
    res[RS_UNCONVERGED] = !conv;
  } else if (a.mode == RELAX_JAM) {
    // The protocol state lives in res (and the saved positions in sx, sy),
    // so a protocol can run over several launches with the same result.
    const JamParams& jp = a.jam;
    R* sx = a.sx + T.v0;
    R* sy = a.sy + T.v0;
    const double* prev = a.res + (size_t)blockIdx.x * RS_COUNT;
    double phi, phi_lo, phi_hi, iters, unconv;
    int stage, status;
    if (a.jam_start) {
      rescale_box(g, T, l, (R)sqrt(A0 / jp.phi0));
      phi = jp.phi0;
      phi_lo = phi_hi = iters = unconv = 0.0;
      stage = 0;
      status = 2;
    } else {
      phi = prev[RS_PHI_TARGET];
      phi_lo = prev[RS_PHI_LO];
      phi_hi = prev[RS_PHI_HI];
      iters = prev[RS_ITER];
      unconv = prev[RS_UNCONVERGED];
      stage = (int)prev[RS_STAGES];
      status = (int)prev[RS_STATUS];
    }
    __syncthreads();  // every thread has read res before thread 0 rewrites it
    for (int budget = a.jam_stages; status == 2 && budget > 0; budget--) {
      bool conv, cleared;
      iters += fire_min(a, T, l, jp.u_jam, conv, cleared);
      unconv += !conv && !cleared;
      stage++;
      // jammed: the minimisation ended (converged or at max_iter) with a
      // vertex-repulsion energy per cell above u_jam
      if (cleared) {
        phi_lo = phi;
        copy_positions(T, l.x, l.y, sx, sy);  // keep the unjammed state
      } else {
        phi_hi = phi;
        if (phi_lo == 0.0) {  // phi0 is already jammed
          status = 1;
          break;
        }
        copy_positions(T, sx, sy, l.x, l.y);  // back to the unjammed state
        T.L = (R)sqrt(A0 / phi_lo);
        phi = phi_lo;
      }
      if (phi_hi > 0.0 && phi_hi - phi_lo < jp.phi_tol) {
        status = 0;
        break;
      }
      if (stage >= jp.max_stages) {
        status = 3;
        break;
      }
      const double next = phi_hi > 0.0 ? 0.5 * (phi_lo + phi_hi) : phi + jp.dphi;
      rescale_box(g, T, l, (R)sqrt(A0 / next));
      phi = next;
    }
    if ((status == 0 || status == 3) && phi != phi_lo) {  // end on the largest unjammed state
      copy_positions(T, sx, sy, l.x, l.y);
      T.L = (R)sqrt(A0 / phi_lo);
      phi = phi_lo;
    }
    res[RS_ITER] = iters;
    res[RS_PHI_LO] = phi_lo;
    res[RS_PHI_HI] = phi_hi;
    res[RS_STAGES] = stage;
    res[RS_UNCONVERGED] = unconv;
    res[RS_STATUS] = status;
    res[RS_PHI_TARGET] = phi;
  }
  if (a.mode != RELAX_EULER) {
    // final state: forces at the final positions, energy (Euler keeps the
    // forces of its last step, as the reference's Forces)
    double U, Ui, fm;
    energy_pass(a, T, l, U, Ui, fm);
    res[RS_FMAX] = fm;
    res[RS_CONVERGED] = fm < a.fire.ftol;
    res[RS_ENERGY] = U;
    res[RS_EINT] = Ui;
  }
  res[RS_L] = (double)T.L;
  res[RS_PHI] = A0 / ((double)T.L * (double)T.L);
  for (int k = tid; k < T.nvt; k += NT) {
    a.x[T.v0 + k] = l.x[k];
    a.y[T.v0 + k] = l.y[k];
  }
  if (tid == 0) {
    for (int i = 0; i < RS_COUNT; i++) a.res[(size_t)blockIdx.x * RS_COUNT + i] = res[i];
    a.tLout[blockIdx.x] = T.L;
  }
}

extern "C" __global__ void __launch_bounds__(1024) k2d_relax_f32(const Relax2D<float> a) { k2d_relax_body(a); }
extern "C" __global__ void __launch_bounds__(1024) k2d_relax_f64(const Relax2D<double> a) { k2d_relax_body(a); }
