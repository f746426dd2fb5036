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
//   k2d_wide_cand   one thread per cell: its candidate cells, ascending;
//   k2d_wide_step   one thread per vertex: all force terms (the crossing
//                   parity by the per-vertex loop), then x + F dt into the
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
  l.ncand = a.ncand + T.c0;
  l.cand = a.cand + (size_t)T.c0 * DPM2D_MAXCAND;
  l.cbits = nullptr;
  l.red = nullptr;
  return l;
}

extern "C" __global__ void __launch_bounds__(256) k2d_wide_cells(const Wide2D a) {
  const int gc = blockIdx.x * blockDim.x + threadIdx.x;
  if (a.counters && gc == 0) atomicAdd(&a.counters[0], 1ull);
  if (gc >= a.C) return;
  const int t = a.cell_tissue[gc];
  const Tissue2DCtx<float> T = tissue_ctx(a.g, t);
  cell_one(a.g, T, wide_view(a, T, t, a.xin, a.yin), gc - T.c0);
}

extern "C" __global__ void __launch_bounds__(256) k2d_wide_cand(const Wide2D a) {
  const int gc = blockIdx.x * blockDim.x + threadIdx.x;
  if (gc >= a.C) return;
  const int t = a.cell_tissue[gc];
  const Tissue2DCtx<float> T = tissue_ctx(a.g, t);
  cand_one(a.g, T, wide_view(a, T, t, a.xin, a.yin), gc - T.c0);
}

extern "C" __global__ void __launch_bounds__(256) k2d_wide_step(const Wide2D a) {
  const int gv = blockIdx.x * blockDim.x + threadIdx.x;
  if (gv >= a.V) return;
  const int gc = cell_of(a.g.cell_vert_begin, a.C, gv);
  const int t = a.cell_tissue[gc];
  const Tissue2DCtx<float> T = tissue_ctx(a.g, t);
  const Lds2D<float> l = wide_view(a, T, t, a.xin, a.yin);
  const int k = gv - T.v0;
  float fx = 0.0f, fy = 0.0f;
  int no = 0, na = 0;
  force_vertex(a.g, T, l, k, fx, fy, &no, a.p_attr ? &na : nullptr, (double*)nullptr, (double*)nullptr,
               (const unsigned*)nullptr, a.p_attr_key ? a.p_attr_key + gv : nullptr);
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
