// The 2D step for tissues beyond one workgroup (float).
//
// k2d_run keeps a whole tissue in one workgroup's LDS; a tissue larger than
// that (the reference's Tissue2D has no size limit: its NDRange covers every
// cell, Tissue2D.cpp:212) runs here over the whole device, three launches
// per step on global arrays laid out as k2d_run's LDS (the same Lds2D view,
// pointing at global memory, and the same per-cell, candidate and per-vertex
// functions, so every decision and every force is k2d_run's):
//   k2d_wide_cells  one thread per cell: strain, orientation, centre of mass,
//                   bounds of the input positions;
//   k2d_wide_cand   one wave per cell: its candidate cells, ascending (the
//                   lanes test 64 cells at a time, a ballot keeps the order);
//   k2d_wide_over   one thread per (vertex, candidate slot): the crossing
//                   parity of the vertex against that candidate, ORed into
//                   a bit per vertex (k2d_run's over_pass over the device;
//                   an OR, so the same bit as the per-vertex loop's);
//   k2d_wide_step   one thread per vertex: all force terms (the crossing
//                   parity from the bits), then x + F dt into the
//                   other position buffer (positions are double-buffered, so
//                   every vertex reads the step's input positions).
#pragma once

struct Wide2D {
  Geom2D<float> g;
  const float* xin;
  const float* yin;
  float* xout;
  float* yout;
  float* fx;  // forces of the step (written with W2_FORCES)
  float* fy;
  const int* cell_tissue;  // [C]
  const int* vbl;          // [C + T]: tissue-local vertex offsets; tissue t's nc + 1 at c0 + t
  float* cellq;            // [8 * C]: strain, orientation, com x, com y, x0, x1, y0, y1
  int* ncand;              // [C]
  short* cand;             // [C * DPM2D_MAXCAND]
  int C, V;
  float dt;
  int flags;
  unsigned long long* counters;  // [4 * DPM_COUNTER_SLOTS] (may be nullptr)
  int* p_overlap;          // diagnostics per vertex (nullptr normally)
  int* p_attr;
  unsigned* p_attr_key;
  unsigned* over_bits;     // [ceil(V / 32)] crossing parity per vertex (k2d_wide_over)
  const int* vcell;        // [V] cell of each vertex
};

enum : int {
  W2_FORCES = 1,  // write the forces
  W2_MOVE = 2,    // write x + F dt
};

// The Lds2D view of tissue t over the global arrays and positions x, y.
__device__ __forceinline__ Lds2D<float> wide_view(const Wide2D& a, const Tissue2DCtx<float>& T, int t, const float* x,
                                                  const float* y) {
  Lds2D<float> l;
  l.x = const_cast<float*>(x) + T.v0;
  l.y = const_cast<float*>(y) + T.v0;
  l.vb = const_cast<int*>(a.vbl) + T.c0 + t;
  float* q = a.cellq + T.c0;
  l.strain = q;
  l.orient = q + a.C;
  l.comx = q + 2 * (size_t)a.C;
  l.comy = q + 3 * (size_t)a.C;
  l.ax0 = q + 4 * (size_t)a.C;
  l.ax1 = q + 5 * (size_t)a.C;
  l.ay0 = q + 6 * (size_t)a.C;
  l.ay1 = q + 7 * (size_t)a.C;
  l.l0c = const_cast<float*>(a.g.l0) + T.c0;
  l.ncand = a.ncand + T.c0;
  l.cand = a.cand + (size_t)T.c0 * DPM2D_MAXCAND;
  l.cbits = nullptr;
  l.red = nullptr;
  return l;
}

extern "C" __global__ void __launch_bounds__(256) k2d_wide_cells(const Wide2D a) {
  const int gc = blockIdx.x * blockDim.x + threadIdx.x;
  // the crossing-parity bits of this step start at zero (a kernel, not a
  // memset: the step is captured in a graph)
  for (int w = gc, nw = (a.V + 31) / 32; w < nw; w += gridDim.x * blockDim.x) a.over_bits[w] = 0u;
  if (a.counters && gc == 0) atomicAdd(&a.counters[0], 1ull);
  if (gc >= a.C) return;
  const int t = a.cell_tissue[gc];
  const Tissue2DCtx<float> T = tissue_ctx(a.g, t);
  cell_one(a.g, T, wide_view(a, T, t, a.xin, a.yin), gc - T.c0);
}

// The list cand_one makes (maybe_pair over the tissue's cells, ascending,
// the count past DPM2D_MAXCAND meaning "scan every cell"), by one wave.
extern "C" __global__ void __launch_bounds__(256) k2d_wide_cand(const Wide2D a) {
  const int lane = threadIdx.x & (DPM_WAVE - 1);
  const int gc = blockIdx.x * (blockDim.x / DPM_WAVE) + (int)(threadIdx.x / DPM_WAVE);
  if (gc >= a.C) return;  // whole wave
  const int t = a.cell_tissue[gc];
  const Tissue2DCtx<float> T = tissue_ctx(a.g, t);
  const Lds2D<float> l = wide_view(a, T, t, a.xin, a.yin);
  const int ci = gc - T.c0;
  int n = 0;
  for (int c0 = 0; c0 < T.nc; c0 += DPM_WAVE) {  // wave-uniform
    const int cj = c0 + lane;
    const bool hit = cj < T.nc && cj != ci && maybe_pair(a.g, T, l, ci, cj);
    const unsigned long long m = __ballot(hit);
    const int slot = n + (int)__popcll(m & ((1ull << lane) - 1ull));
    if (hit && slot < DPM2D_MAXCAND) l.cand[ci * DPM2D_MAXCAND + slot] = (short)cj;
    n += (int)__popcll(m);
  }
  if (lane == 0) l.ncand[ci] = n;
}

// Threads (slot q, vertex gv), q-major: the lanes of a wave take consecutive
// vertices (mostly one cell's, with one candidate list) and the same slot.
// A cell with more than DPM2D_MAXCAND candidates scans every cell: slot q
// then takes cells q, q + DPM2D_MAXCAND, ....  Launched with repulsion only.
extern "C" __global__ void __launch_bounds__(256) k2d_wide_over(const Wide2D a) {
  const long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= (long long)a.V * DPM2D_MAXCAND) return;
  const int q = (int)(w / a.V), gv = (int)(w % a.V);
  const int gc = a.vcell[gv];
  const int t = a.cell_tissue[gc];
  const Tissue2DCtx<float> T = tissue_ctx(a.g, t);
  const Lds2D<float> l = wide_view(a, T, t, a.xin, a.yin);
  const int ci = gc - T.c0, k = gv - T.v0;
  const int ncd = l.ncand[ci];
  const float px = l.x[k], py = l.y[k];
  bool odd = false;
  if (ncd > DPM2D_MAXCAND) {
    for (int cj = q; cj < T.nc && !odd; cj += DPM2D_MAXCAND)
      if (cj != ci && !outside_box_even(l, cj, px, py, T.pbc, T.L)) odd = crossing_odd(l, cj, px, py, T.pbc, T.L);
  } else if (q < ncd) {
    const int cj = l.cand[ci * DPM2D_MAXCAND + q];
    if (cj != ci && !outside_box_even(l, cj, px, py, T.pbc, T.L)) odd = crossing_odd(l, cj, px, py, T.pbc, T.L);
  }
  if (odd) atomicOr(&a.over_bits[gv >> 5], 1u << (gv & 31));
}

extern "C" __global__ void __launch_bounds__(256) k2d_wide_step(const Wide2D a) {
  const int gv = blockIdx.x * blockDim.x + threadIdx.x;
  if (gv >= a.V) return;
  const int gc = a.vcell[gv];
  const int t = a.cell_tissue[gc];
  const Tissue2DCtx<float> T = tissue_ctx(a.g, t);
  const Lds2D<float> l = wide_view(a, T, t, a.xin, a.yin);
  const int k = gv - T.v0;
  float fx = 0.0f, fy = 0.0f;
  int no = 0, na = 0;
  const int over = (a.g.mask & M2_REPULSION) ? (int)((a.over_bits[gv >> 5] >> (gv & 31)) & 1u) : -1;
  force_vertex(a.g, T, l, k, fx, fy, &no, a.p_attr ? &na : nullptr, (double*)nullptr, (double*)nullptr,
               (const unsigned*)nullptr, a.p_attr_key ? a.p_attr_key + gv : nullptr, over);
  if (a.p_overlap) a.p_overlap[gv] = no;
  if (a.p_attr) a.p_attr[gv] = na;
  if (a.flags & W2_FORCES) {
    a.fx[gv] = fx;
    a.fy[gv] = fy;
  }
  if (a.flags & W2_MOVE) {
    a.xout[gv] = fmaf(fx, a.dt, a.xin[gv]);
    a.yout[gv] = fmaf(fy, a.dt, a.yin[gv]);
  }
  if (a.counters && no) atomicAdd(&a.counters[4 * (gv % DPM_COUNTER_SLOTS) + 2], 1ull);
}
