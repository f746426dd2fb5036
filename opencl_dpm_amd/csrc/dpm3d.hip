// 3D DPM timestep kernels for gfx950.
//
// One workgroup owns one cell for a whole step (grid = cells of all tissues,
// workgroups spread over the 8 XCDs so that neighbouring cells share an L2).
// The cell's vertices and the mesh's face table live in LDS for the step.
// Every force term of the reference's per-face kernels is evaluated once per
// face into LDS and *gathered* by each vertex over its incident faces in a
// fixed order, so no float atomics are used and a step is bit-reproducible
// run to run (the reference scatters with a compare-and-swap float add,
// Cell3D_Kernel.cl:10-32, whose order varies).
//
// Positions are double-buffered (read set -> write set) and the per-cell
// statistics the next step needs (centre of mass, bounding radius, volume,
// longest edge) are reduced from the new positions in the same launch, so a
// step is ONE kernel launch reading and writing each vertex once.
//
// The kernels are templates on the real type R: R = float is the production
// form (float forces, double per-cell sums); R = double evaluates the same
// formulas in double (float thresholds promoted, as the FP64 restatement of
// the oracle does) and is the full-FP64 reference run of the precision form.
//
// Reference semantics followed (Cell3D_Kernel.cl):
//   VolumeForceUpdate       :66-112   (volume of the current positions)
//   SurfaceAreaForceUpdate  :114-177
//   StickToSurface          :180-247  (optionally with a stiffness gradient)
//   RepellingForces         :251-310  (winding number, culled exactly)
//   AllVertAttraction       :313-364  (not launched by the reference host)
//   EulerPosition           :371-381
// plus graph-Laplacian bending, which has no reference kernel.
#include "dpm_device.h"

using namespace dpm;

#define DPM3D_MAXDEG 8
#define DPM3D_MAXPART 4

enum : unsigned {
  M_VOLUME = 1u,
  M_SURFACE = 2u,
  M_SUBSTRATE = 4u,
  M_CONTACT = 8u,
  M_ADHESION = 16u,
  M_BENDING = 32u,
};
enum : int {
  F_WRITE_FORCES = 1,
  F_MOVE = 2,
};

// Per-cell statistics of a step's input positions (written by the previous
// step's epilogue or k3d_stats): centre of mass, bounding radius, longest
// edge, target edge length l0, and 1 / max(l0, 1e-12) in double (the
// adhesion term's, divided once per cell instead of once per pair).  Eight
// reals either way (32 bytes in float, 64 in double).
template <typename R>
struct CellStat;
template <>
struct CellStat<float> {
  float cx, cy, cz, r;
  float em, l0;
  double inv_l0;
};
template <>
struct CellStat<double> {
  double cx, cy, cz, r;
  double em, l0;
  double inv_l0, pad;
};
template <typename R>
__host__ __device__ inline CellStat<R> make_cell_stat(R cx, R cy, R cz, R r, R em, R l0) {
  CellStat<R> s{};
  s.cx = cx;
  s.cy = cy;
  s.cz = cz;
  s.r = r;
  s.em = em;
  s.l0 = l0;
  s.inv_l0 = 1.0 / (double)(l0 > R(1e-12f) ? l0 : R(1e-12f));
  return s;
}

template <typename R>
struct alignas(4 * sizeof(R)) Q4 {
  R x, y, z, w;
};

template <typename R>
struct Step3D {
  const R* xin;
  const R* yin;
  const R* zin;
  R* xout;
  R* yout;
  R* zout;
  R* fx;
  R* fy;
  R* fz;
  const CellStat<R>* st_in;
  const double* vol_in;
  const R* em_in;  // per cell: longest edge
  CellStat<R>* st_out;
  double* vol_out;
  R* em_out;
  const float* Kv;
  const float* Ka;
  const float* Ks;
  const float* Kb;
  const float* v0;
  const float* a0;
  const int* cell_tissue;
  const int* tissue_begin;
  const float* tL;
  const int* tpbc;
  const float* tKre;
  const float* tKat;
  const float* tksg;
  const int* faces;         // [nf*3]
  const int* vf;            // [nv*MAXDEG] incident (face*3 + corner), ascending, -1 pad
  const int* adj;           // [nv*MAXDEG] neighbour vertices (ring order if closed), -1 pad
  int nv, nf, ncells;
  int closed;  // every edge on two faces, every vertex's faces one oriented cycle
  // the wave form's workgroup-shared topology, byte for byte its LDS image
  // (packed faces, 16-bit incidence and neighbour rows), in 16-byte words
  const uint4* wtopo;
  int wtopo_n16;
  int ww;  // waves that evaluate winding numbers (LDS tables)
  R dt;
  unsigned mask;
  int flags;
  unsigned long long* counters;  // [4 * DPM_COUNTER_SLOTS]: steps, windings, contacts, candidates
  // diagnostics (nullptr when not probing)
  unsigned char* p_contact;
  float* p_winding;
  unsigned char* p_down;
  int* p_adh;
  // per vertex: key sums (vertex_key) of the adhesion partners its own loop
  // accepts, then of the partners whose loop accepts it
  unsigned* p_adh_hash;
  // per vertex: number of contact partner cells, key sum of their indices
  // within the tissue (cell_key)
  unsigned* p_contact_key;
  int ncmax;
  double* energy;  // k3d_energy output, per cell
  R* w_out;        // per vertex: sum of |w| over its contact partners (nullptr: not recorded)
  // contact partners per vertex for the adjoint (nullptr: not recorded):
  // up to DPM3D_MAXPART of them, cell (-1 pad), winding number, shift
  int* rec_cell;      // [nverts * MAXPART]
  R* rec_w;           // [nverts * MAXPART]
  R* rec_shift;       // [nverts * MAXPART * 3]
  int* rec_overflow;  // set when a vertex has more partners
  // compensated positions (nullptr: plain update): the position of a vertex
  // is x + xlo, xlo the rounding residual carried from step to step
  R* xlo;
  R* ylo;
  R* zlo;
  // cell list of this step (nullptr: scan every cell of the tissue)
  const int* bin_start;  // per tissue: [nbuck + 1] bucket offsets at tbin_hoff
  const int* bin_list;   // cells of each tissue ordered by bucket
  const int* tbin_nb;    // per tissue: bins per axis (0: scan)
  const int* tbin_nbuck;
  const int* tbin_hoff;
};

// a + b = s + e exactly (Knuth's TwoSum, any magnitudes).
template <typename R>
__device__ __forceinline__ R two_sum(R a, R b, R& e) {
  const R s = a + b;
  const R bb = s - a;
  e = (a - (s - bb)) + (b - bb);
  return s;
}

// Target edge length l0 = sqrt(4 a0)/sqrt(3) of a cell (Tissue3D.cpp:177),
// in the real type (for float, the same correctly rounded operations as the
// reference host's sqrt on float).
template <typename R>
__device__ __forceinline__ R cell_l0(const float* a0, int c) {
  return sqrt(R(4) * (R)a0[c]) / sqrt(R(3));
}

// Periodic shift of cell j toward cell i along one axis, L round(d / L)
// with d = C_i - C_j (Cell3D_Kernel.cl:279-281).  The quotient is taken as
// a product with 1/L (within two ulp of the quotient); only where that
// product lies within 1e-5 (relative) of a half-integer can the two round
// differently, and there the division itself decides.
template <typename R>
__device__ __forceinline__ R image_shift(R d, R L, R invL) {
  const R q = d * invL;
  R n = round(q);
  if (fabs(fabs(q - n) - R(0.5f)) < R(1e-5f) * (R(1) + fabs(q))) n = round(d / L);
  return L * n;
}

// Culling reach added to a cell's bounding radius for the winding-number
// test.  Every face of cell j lies inside its bounding sphere R_j, so a
// point p outside R_j + m_j is at least m_j from every face and edge of j,
// and outside j's convex hull: its exact winding number is 0.  An edge of
// length <= e seen from distance >= d subtends at most 2 atan(e / 2d); with
// d >= 0.4 e that is <= 102.7 degrees, so every pair of the unit vectors
// u, v, w toward a face's corners has a dot product >= -0.2195 and the
// reference's denominator 1 + u.v + v.w + w.u (Cell3D_Kernel.cl:293) is
// >= 0.34: no face is skipped (denom < 1e-8), and the float sum of the
// face angles is rounding noise about a true 0, far below the 1e-6 pair
// threshold (Cell3D_Kernel.cl:303) — the pair is absent in the reference
// arithmetic too.  m_j = 0.4 e_max(j) (with slack for float rounding).
template <typename R>
__device__ __forceinline__ R contact_margin(R Rj, R emax_j) {
  return emax_j * R(0.4004f) + R(1e-6f) * Rj;
}

// Exact outside test ahead of a winding number, evaluated by one whole wave
// (every lane calls it with the same arguments).  With d = p - C_j and
// h = max_v (x_v - c_j).d over cell j's vertices (x_v unshifted, c_j its
// unshifted centre), every point of j's convex hull lies in
// {x : (x - C_j).d <= h}, so p is at least (|d|^2 - h) / |d| from every
// face of j.  Where that distance is at least the culling margin m_j, the
// argument of contact_margin holds (no face is skipped, the point is outside
// the hull): the winding number is 0 up to rounding and the contact test
// rejects the point.  The margin is raised by 1e-4 (|d| + R_j) and by 1e-6
// of the coordinates' magnitude (c_j and the shifted C_j) for the rounding of
// these float sums.
template <typename R>
__device__ __forceinline__ bool outside_by_margin(const R* xs, const R* ys, const R* zs, int nv, R cx, R cy,
                                                  R cz, V3<R> Cj, V3<R> d, R m, R Rj) {
  const int lane = threadIdx.x & (DPM_WAVE - 1);
  R h = R(-3.0e38f);
  for (int v = lane; v < nv; v += DPM_WAVE)
    h = fmax(h, ((xs[v] - cx) * d.x + (ys[v] - cy) * d.y) + (zs[v] - cz) * d.z);
  h = wave_max_all(h);
  const R dn = sqrt(vdot(d, d));
  const R M = ((fabs(cx) + fabs(cy)) + fabs(cz)) + ((fabs(Cj.x) + fabs(Cj.y)) + fabs(Cj.z));
  return dn > R(0) && dn * dn - h >= ((m + R(1e-4f) * (dn + Rj)) + R(1e-6f) * M) * dn;
}

// Edge-spring factor (|d|/l0 - 1)/|d| of an edge vector with |d|^2 = d2
// (Cell3D_Kernel.cl:140-176): in float 1/l0 - rsqrt(d2) by the hardware
// reciprocal square root (about 1 ulp), in double as written.
template <typename R>
__device__ __forceinline__ R spring_factor(R d2, R inv_l0) {
  if constexpr (sizeof(R) == 4) {
    return inv_l0 - __builtin_amdgcn_rsqf(d2);
  } else {
    const R len = sqrt(d2);
    return (len * inv_l0 - R(1)) / len;
  }
}

// Facing-down decision of a face with normal n = (P1 - P0) x (P2 - P0):
// normalize(n).z < -0.1 (Cell3D_Kernel.cl:206-214), decided from n_z and |n|^2
// where n_z^2 and 0.01 |n|^2 are clear of each other by 1e-5 (relative), else
// by the normalised vector itself (the reference's operation order).
template <typename R>
__device__ __forceinline__ bool facing_down(V3<R> n) {
  const R q = n.z * n.z, t = (R(0.1f) * R(0.1f)) * vdot(n, n);
  if (n.z >= R(0) || q < t * R(0.99999f)) return false;
  if (q > t * R(1.00001f)) return true;
  return vnormalize(n).z < R(-0.1f);
}

// LDS layout (byte offsets), shared by host (sizing) and device (carving).
// The per-(face, corner) shape forces are dead once every vertex gathered
// them, so they share their bytes with the tables of the later passes
// (bending, neighbour cell, winding directions).
struct Lds3DLayout {
  size_t own, faces, clist, plist, glist, gath, cnt, red, cst, nbr, bend, wnum, utab, fc, total;
};
// Statistics of the first kCandStats candidates of a scan chunk are kept in
// LDS (the rest are read from global memory).
constexpr int kCandStats = 64;
__host__ __device__ inline Lds3DLayout lds3d_layout(int nv, int nf, int nt, int ww, int rsize) {
  Lds3DLayout L;
  const size_t rs = (size_t)rsize;
  size_t o = 0;
  L.own = o;     o += 3u * rs * (size_t)nv;
  o = (o + 7u) & ~(size_t)7u;
  L.faces = o;   o += 8u * (size_t)nf;  // (i0 | i1 << 16, i2) per face
  o = (o + 3u) & ~(size_t)3u;
  L.clist = o;   o += 4u * (size_t)nt;
  L.plist = o;   o += 8u * (size_t)nt;  // two lists of nt
  L.glist = o;   o += 4u * (size_t)nt;  // candidate cells gathered from the cell list
  L.gath = o;    o += 4u * 96u;         // 27 buckets: ids, starts, prefix
  L.cnt = o;     o += 4u * 48u;
  o = (o + 7u) & ~(size_t)7u;
  L.red = o;     o += 8u * 192u;  // three 64-double regions (cell_stats alternates them)
  o = (o + 31u) & ~(size_t)31u;
  L.cst = o;     o += 8u * rs * (size_t)kCandStats;  // CellStat<R> of the candidates
  o = (o + 31u) & ~(size_t)31u;
  L.fc = o;  // union: shape pass | later passes
  const size_t fc_end = o + 9u * rs * (size_t)nf;
  L.nbr = o;     o += 3u * rs * (size_t)nv;
  L.bend = o;    o += 3u * rs * (size_t)nv;
  L.wnum = o;    o += rs * (size_t)nv;
  o = (o + 31u) & ~(size_t)31u;
  L.utab = o;    o += 4u * rs * (size_t)nv * (size_t)ww;
  // the adhesion pass's per-(slice, vertex) partial sums share the winding
  // directions' bytes (the two passes are separated by barriers)
  const size_t adh_end = L.utab + 4u * rs * (size_t)nt;
  if (o < adh_end) o = adh_end;
  L.total = o > fc_end ? o : fc_end;
  return L;
}

template <typename R>
struct Lds3D {
  R *sx, *sy, *sz, *nx, *ny, *nz, *sb, *wb, *fcx, *fcy, *fcz;
  Q4<R>* utab;
  CellStat<R>* cst;
  uint2* sf;
  int *clist, *plist, *glist, *gath, *cnt;
  double* red;
};
template <typename R>
__device__ __forceinline__ Lds3D<R> carve3d(unsigned char* smem, int nv, int nf, int nt, int ww) {
  const Lds3DLayout L = lds3d_layout(nv, nf, nt, ww, (int)sizeof(R));
  Lds3D<R> l;
  R* o = reinterpret_cast<R*>(smem + L.own);
  l.sx = o;
  l.sy = o + nv;
  l.sz = o + 2 * nv;
  R* n = reinterpret_cast<R*>(smem + L.nbr);
  l.nx = n;
  l.ny = n + nv;
  l.nz = n + 2 * nv;
  l.sb = reinterpret_cast<R*>(smem + L.bend);
  l.wb = reinterpret_cast<R*>(smem + L.wnum);
  l.sf = reinterpret_cast<uint2*>(smem + L.faces);
  l.clist = reinterpret_cast<int*>(smem + L.clist);
  l.plist = reinterpret_cast<int*>(smem + L.plist);
  l.glist = reinterpret_cast<int*>(smem + L.glist);
  l.gath = reinterpret_cast<int*>(smem + L.gath);
  l.cnt = reinterpret_cast<int*>(smem + L.cnt);
  l.utab = reinterpret_cast<Q4<R>*>(smem + L.utab);
  l.fcx = reinterpret_cast<R*>(smem + L.fc);
  l.fcy = l.fcx + 3 * nf;
  l.fcz = l.fcy + 3 * nf;
  l.red = reinterpret_cast<double*>(smem + L.red);
  l.cst = reinterpret_cast<CellStat<R>*>(smem + L.cst);
  return l;
}

// Orders this wave's LDS writes before its later LDS reads (the LDS unit
// executes one wave's operations in order; this keeps the compiler from
// reordering them and waits for outstanding LDS traffic).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

template <typename R>
__device__ __forceinline__ V3<R> lds_v(const R* x, const R* y, const R* z, int i) {
  return {x[i], y[i], z[i]};
}

template <typename R>
__device__ __forceinline__ void face_of(const Lds3D<R>& l, int f, int& i0, int& i1, int& i2) {
  const uint2 w = l.sf[f];
  i0 = (int)(w.x & 0xffffu);
  i1 = (int)(w.x >> 16);
  i2 = (int)w.y;
}

// Centre of mass (double sums), bounding radius, volume and longest edge of
// the cell whose vertices are in l.sx/sy/sz.  The volume is summed over
// faces relative to the centre of mass (translation-invariant for a closed
// surface), per-face terms in R, the sum in double — the reference sums
// absolute-coordinate triple products in float (getVolume,
// Cell3D_Kernel.cl:46-64), which cancels catastrophically far from the
// origin.
// one_each: a thread per vertex, vertex tid in pv (cell_stats_own).
template <typename R>
__device__ void cell_stats_impl(const Lds3D<R>& l, int nv, int nf, CellStat<R>* st_out,
                                double* vol_out, R* em_out, int ci, R l0c, bool one_each, bool has_v, V3<R> pv) {
  const int tid = threadIdx.x, NT = blockDim.x;
  double px = 0.0, py = 0.0, pz = 0.0;
  if (one_each) {
    if (has_v) {
      px = (double)pv.x;
      py = (double)pv.y;
      pz = (double)pv.z;
    }
  } else {
    for (int v = tid; v < nv; v += NT) {
      px += (double)l.sx[v];
      py += (double)l.sy[v];
      pz += (double)l.sz[v];
    }
  }
  block_sum3_1b(px, py, pz, l.red + 64);  // the two reductions alternate scratch regions
  const V3<R> C = {(R)(px / nv), (R)(py / nv), (R)(pz / nv)};
  R r = R(0);
  if (one_each) {
    if (has_v) {
      const V3<R> d = vsub(pv, C);
      r = sqrt(vdot(d, d));
    }
  } else {
    for (int v = tid; v < nv; v += NT) {
      const V3<R> d = vsub(lds_v(l.sx, l.sy, l.sz, v), C);
      r = fmax(r, sqrt(vdot(d, d)));
    }
  }
  double vol = 0.0;
  R e2 = R(0);
  for (int f = tid; f < nf; f += NT) {
    int i0, i1, i2;
    face_of(l, f, i0, i1, i2);
    const V3<R> P0 = vsub(lds_v(l.sx, l.sy, l.sz, i0), C);
    const V3<R> P1 = vsub(lds_v(l.sx, l.sy, l.sz, i1), C);
    const V3<R> P2 = vsub(lds_v(l.sx, l.sy, l.sz, i2), C);
    vol += (double)vdot(vcross(P0, P1), P2);
    const V3<R> e0 = vsub(P1, P0), e1 = vsub(P2, P1), ee = vsub(P0, P2);
    e2 = fmax(e2, fmax(vdot(e0, e0), fmax(vdot(e1, e1), vdot(ee, ee))));
  }
  {  // the radius, volume and longest edge in one reduction (max of
     // non-negative floats is exact in double)
    double rd = (double)r, e2d = (double)e2;
    block_max_sum_max_1b<sizeof(R) == 4>(rd, vol, e2d, l.red + 128);
    r = (R)rd;
    e2 = (R)e2d;
  }
  if (tid == 0) {
    // Slack of a few ulp so that the culling tests are never tighter than the
    // largest distance a vertex can have from C, or the longest edge.
    const R em = sqrt(e2) * R(1.000001f);
    st_out[ci] = make_cell_stat<R>(C.x, C.y, C.z, r * R(1.0000005f) + R(1e-30f), em, l0c);
    vol_out[ci] = fabs(vol / 6.0);
    em_out[ci] = em;
  }
}
template <typename R>
__device__ void cell_stats(const Lds3D<R>& l, int nv, int nf, CellStat<R>* st_out, double* vol_out, R* em_out,
                           int ci, R l0c) {
  cell_stats_impl(l, nv, nf, st_out, vol_out, em_out, ci, l0c, false, false, V3<R>{R(0), R(0), R(0)});
}
// The same from a workgroup with a thread per vertex (NT >= nv): thread v
// holds vertex v (has_v) in registers, and the centre of mass and radius are
// taken from them (the sums and maxima of cell_stats, bit for bit); the LDS
// positions are read only by the face loop, after the first reduction's
// barrier, so they may be written just before the call.
template <typename R>
__device__ void cell_stats_own(const Lds3D<R>& l, int nv, int nf, CellStat<R>* st_out, double* vol_out,
                               R* em_out, int ci, R l0c, bool has_v, V3<R> pv) {
  cell_stats_impl(l, nv, nf, st_out, vol_out, em_out, ci, l0c, true, has_v, pv);
}

template <typename R>
__device__ __forceinline__ void stage_topology(const Step3D<R>& a, const Lds3D<R>& l) {
  const int tid = threadIdx.x, NT = blockDim.x;
  for (int f = tid; f < a.nf; f += NT)
    l.sf[f] = make_uint2((unsigned)a.faces[3 * f] | ((unsigned)a.faces[3 * f + 1] << 16),
                         (unsigned)a.faces[3 * f + 2]);
}

template <typename R>
__device__ __forceinline__ void load_own(const Step3D<R>& a, const Lds3D<R>& l, size_t base) {
  for (int v = threadIdx.x; v < a.nv; v += blockDim.x) {
    l.sx[v] = a.xin[base + v];
    l.sy[v] = a.yin[base + v];
    l.sz[v] = a.zin[base + v];
  }
}

// The reference's adhesion decision sqrt(d2) < 2 l0 && sqrt(d2) > 1e-12
// (Cell3D_Kernel.cl:348-353), for float without the square root: with m the
// midpoint between a positive float c and the float below it, RN(sqrt(x)) < c
// exactly when x < m^2, and RN(sqrt(x)) > c exactly when x > m'^2 (m' the
// midpoint with the float above).  m has at most 25 significant bits, so m^2
// is exact in double, and sqrt(x) never equals m for a float x (m^2 would
// need at least 49 significant bits), so no rounding tie arises.  Double
// steps take the square root.
template <typename R>
struct AdhCut {
  R c;
  double hi, lo;
  __device__ __forceinline__ explicit AdhCut(R two_l0) : c(two_l0), hi(0.0), lo(0.0) {
    if constexpr (sizeof(R) == 4) {
      const float below = __int_as_float(__float_as_int((float)two_l0) - 1);
      const float tiny = 1e-12f, above = __int_as_float(__float_as_int(tiny) + 1);
      const double m = 0.5 * ((double)below + (double)two_l0);
      const double n = 0.5 * ((double)tiny + (double)above);
      hi = m * m;
      lo = n * n;
    }
  }
  __device__ __forceinline__ bool operator()(R d2) const {
    if constexpr (sizeof(R) == 4) {
      const double x = (double)d2;
      return x < hi && x > lo;
    } else {
      const R d = sqrt(d2);
      return d < c && d > R(1e-12f);
    }
  }
};

// Kat/2 (d/L0 - 1) delta/d with delta = (b - a) - L n, the adhesion term of
// Cell3D_Kernel.cl:354-357, evaluated in double (the step kernels call it
// for pairs across the periodic boundary; a float step's pairs without an
// image take adhesion_term_near).  Float steps take it as
// Kat/2 (1/L0 - 1/d) delta with 1/d from the hardware reciprocal square root
// refined by two Newton steps (each squares the relative error) and kh =
// Kat/2, inv_l0 = 1/L0 taken once per candidate pair — the same double value
// to rounding, a third of the instructions.  Double steps keep the square
// root and divisions (their registers are the tighter constraint there).
template <typename R>
__device__ __forceinline__ V3<R> adhesion_term(V3<R> b, V3<R> a, V3<R> n, R L, R Kat, R L0, double kh,
                                               double inv_l0) {
  const double Ld = (double)L;
  const double dx = ((double)b.x - (double)a.x) - Ld * (double)n.x;
  const double dy = ((double)b.y - (double)a.y) - Ld * (double)n.y;
  const double dz = ((double)b.z - (double)a.z) - Ld * (double)n.z;
  double k;
  if constexpr (sizeof(R) == 4) {
    const double d2 = dx * dx + dy * dy + dz * dz;
    double y = __builtin_amdgcn_rsq(d2);
    y = y * (1.5 - (0.5 * d2) * (y * y));
    y = y * (1.5 - (0.5 * d2) * (y * y));
    k = kh * (inv_l0 - y);
  } else {
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    k = (double)Kat * 0.5 * (d / (double)L0 - 1.0) / d;
  }
  return {(R)(k * dx), (R)(k * dy), (R)(k * dz)};
}

// A float step's adhesion term on a pair without a periodic image, added to
// F: Kat/2 (1/L0 - 1/d) delta is subtracted, with delta = pj - p the float
// difference of two nearby positions (relative error 2^-24) and y = 1/d from
// the hardware reciprocal square root, for relative error near 1e-7 (the
// partner's term of the same pair, Kat/2 (1/L0j - 1/d) (p - pj), is the same
// call with its own 1/L0).  Force arithmetic only, fused.
__device__ __forceinline__ V3<float> adhesion_term_near(V3<float> delta, float y, float kh, float inv_l0,
                                                        V3<float> F) {
  const float k = kh * (inv_l0 - y);
  return vmadd_f(delta, -k, F);
}

// Substrate force on one corner of a facing-down face
// (Cell3D_Kernel.cl:229-245): a z < 0 penalty and the spreading push away
// from the centre of mass within 2 l0 of the plane.  The decisions (z < 0,
// |z| < 2 l0) are the reference's; the float force takes its distance to the
// plane as |z| (the correctly rounded sqrt(z*z) is |z|), |z|/l0 as |z| times
// sub_inv_l0 = 1/l0 in double and the unit vector by the hardware reciprocal
// square root (force arithmetic only, relative error near 1e-7).
template <typename R>
__device__ __forceinline__ V3<R> substrate_corner(V3<R> pos, V3<R> C, R Ks, R ksg, R L, R l0, double sub_inv_l0) {
  V3<R> c = {R(0), R(0), R(0)};
  R ks = Ks;
  if (ksg != R(0)) {
    const R xw = pos.x - L * floor(pos.x / L);
    ks = Ks * (R(1) + ksg * xw / L);
  }
  if (pos.z < R(0)) c.z += ks * fabs(pos.z);
  const R height = fabs(pos.z);
  if (height < l0 * R(2)) {
    const V3<R> c2v = {pos.x - C.x, pos.y - C.y, R(0) - C.z};
    if constexpr (sizeof(R) == 4) {
      const R ftmp = (R)((double)ks * (1.0 - (double)height * sub_inv_l0));
      const R n2 = vdot(c2v, c2v);
      const R s = n2 > R(0) ? __builtin_amdgcn_rsqf(n2) * ftmp : R(0);
      c = vmadd_f(c2v, s, c);
    } else {
      const R dz = pos.z - R(0);
      const R dist = sqrt((R(0) * R(0) + R(0) * R(0)) + dz * dz);
      const R ftmp = (R)((double)ks * (1.0 - (double)(dist / l0)));
      c = vadd(c, vscale(vnormalize(c2v), ftmp));
    }
  }
  return c;
}

// ---------------------------------------------------------------- cell list
//
// Per step and tissue, the cells' centres of mass are binned on a periodic
// grid of nb^3 bins of width L/nb >= the largest candidate reach of the
// step, bound from the current radii, longest edges and l0:
//   2 R_max + max(contact margin, 2 l0_max), with the step's own slack.
// Any candidate pair then lies in neighbouring bins (minimum image), so the
// 27 buckets around a cell hold every candidate; the exact predicate of the
// scan decides, on the candidates sorted by cell index, so the step is
// bit-identical to scanning every cell.  Bins are hashed into nbuck buckets
// (a power of two; collisions only add candidates).  Tissues whose reach is
// not small against L (nb < 3, or the all-images adhesion rule) or without
// periodic boundaries scan every cell.
__device__ __forceinline__ unsigned bin_bucket(int bx, int by, int bz, int nb, int nbuck) {
  const unsigned key = (unsigned)bx + (unsigned)nb * ((unsigned)by + (unsigned)nb * (unsigned)bz);
  if ((unsigned long long)nb * nb * nb <= (unsigned long long)nbuck) return key;
  unsigned h = key * 2654435761u;
  h ^= h >> 16;
  return h & (unsigned)(nbuck - 1);
}

// The distinct buckets among the 27 around a cell, ascending, with their
// cells' starts (gst[u]) and the prefix of their sizes (gof[u]; gof[nu] =
// the total), from one whole wave: lane l < 27 brings bucket id g (the
// other lanes' g is not read).  Hashed buckets may repeat an id; a repeat is
// kept at its first lane only.  Every lane gets nu.  Replaces a serial sort
// on one lane: 27 uniform rounds of lane reads instead of a chain of
// dependent LDS accesses.
__device__ __forceinline__ int distinct_buckets27(int g, const int* bs, int* gst, int* gof) {
  const int lane = threadIdx.x & (DPM_WAVE - 1);
  const bool in = lane < 27;
  bool dup = false;
#pragma unroll
  for (int k = 0; k < 27; k++) dup = dup || (k < lane && __builtin_amdgcn_readlane(g, k) == g);
  const bool keep = in && !dup;
  const unsigned long long km = __ballot(keep);
  const int start = keep ? bs[g] : 0;
  const int cnt = keep ? bs[g + 1] - start : 0;
  int rank = 0, off = 0, total = 0;
#pragma unroll
  for (int k = 0; k < 27; k++) {
    const int gk = __builtin_amdgcn_readlane(g, k), ck = __builtin_amdgcn_readlane(cnt, k);
    const bool below = ((km >> k) & 1ull) && gk < g;
    rank += below ? 1 : 0;
    off += below ? ck : 0;
    total += ck;
  }
  if (keep) {
    gst[rank] = start;
    gof[rank] = off;
  }
  const int nu = (int)__popcll(km);
  if (lane == 0) gof[nu] = total;
  return nu;
}

template <typename R>
__device__ __forceinline__ int bin_coord(R x, R L, R wb, int nb) {
  const R xw = x - L * floor(x / L);
  const int b = (int)floor(xw / wb);
  return b < 0 ? 0 : (b >= nb ? nb - 1 : b);
}

template <typename R>
struct Bin3D {
  const int* tissue_begin;
  const CellStat<R>* st_in;
  const R* em_in;
  const float* a0;
  const float* tL;
  const int* tpbc;
  const float* tKre;
  const float* tKat;
  unsigned mask;
  int* tbin_nb;
  const int* tbin_nbuck;
  const int* tbin_hoff;
  int* bin_start;
  int* bin_list;
  float wscale;  // bin width over the largest candidate reach (>= 1)
};

// One workgroup per tissue; LDS: nbuck counters + bucket and rank per cell.
template <typename R>
__device__ void k3d_bin_body(const Bin3D<R>& a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int t = blockIdx.x, tid = threadIdx.x, NT = blockDim.x;
  const int cb = a.tissue_begin[t], ce = a.tissue_begin[t + 1], nc = ce - cb;
  const int nbuck = a.tbin_nbuck[t];
  int* cnt = reinterpret_cast<int*>(smem);
  int* cbk = cnt + nbuck + 1;
  int* crk = cbk + nc;
  double* red = reinterpret_cast<double*>(smem + (((size_t)(nbuck + 1 + 2 * nc) * 4 + 7) & ~(size_t)7));
  const R L = (R)a.tL[t];
  const bool do_contact = (a.mask & M_CONTACT) && a.tKre[t] != 0.0f;
  const bool do_adh = (a.mask & M_ADHESION) && a.tKat[t] != 0.0f;
  R rmax = R(0), emax = R(0), l0max = R(0);
  for (int k = tid; k < nc; k += NT) {
    rmax = fmax(rmax, a.st_in[cb + k].r);
    emax = fmax(emax, a.em_in[cb + k]);
    l0max = fmax(l0max, fmax(cell_l0<R>(a.a0, cb + k), R(1e-12f)));
  }
  rmax = block_max(rmax, reinterpret_cast<R*>(red));
  emax = block_max(emax, reinterpret_cast<R*>(red));
  l0max = block_max(l0max, reinterpret_cast<R*>(red));
  R reach = R(0);
  if (do_contact) reach = R(2) * rmax + contact_margin(rmax, emax);
  bool all_images = false;
  if (do_adh) {
    const R r2 = R(2) * rmax + R(2) * l0max;
    reach = fmax(reach, r2);
    all_images = R(2) * r2 >= R(0.999f) * L;
  }
  reach = (reach * R(1.0001f) + R(1e-6f)) * R(1.001f);  // the step's slack, and more
  reach *= (R)a.wscale;  // wider bins: fewer, fuller buckets (any width >= the reach is exact)
  int nb = (int)floor(L / reach);
  if (!a.tpbc[t] || all_images || nb < 3 || !(do_contact || do_adh) || nbuck == 0) nb = 0;
  if (nb > 1024) nb = 1024;
  if (tid == 0) a.tbin_nb[t] = nb;
  if (nb == 0) return;  // block-uniform
  const R wb = L / (R)nb;
  for (int b = tid; b <= nbuck; b += NT) cnt[b] = 0;
  __syncthreads();
  for (int k = tid; k < nc; k += NT) {
    const CellStat<R> s = a.st_in[cb + k];
    const unsigned b = bin_bucket(bin_coord(s.cx, L, wb, nb), bin_coord(s.cy, L, wb, nb),
                                  bin_coord(s.cz, L, wb, nb), nb, nbuck);
    cbk[k] = (int)b;
    crk[k] = atomicAdd(&cnt[b], 1);
  }
  __syncthreads();
  if (tid == 0) {  // exclusive scan (nbuck <= 8192)
    int run = 0;
    for (int b = 0; b < nbuck; b++) {
      const int c = cnt[b];
      cnt[b] = run;
      run += c;
    }
    cnt[nbuck] = run;
  }
  __syncthreads();
  int* bs = a.bin_start + a.tbin_hoff[t];
  for (int b = tid; b <= nbuck; b += NT) bs[b] = cnt[b];
  for (int k = tid; k < nc; k += NT) a.bin_list[cb + cnt[cbk[k]] + crk[k]] = cb + k;
}

template <typename R>
__device__ void k3d_stats_body(const Step3D<R>& a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int ci = xcd_remap(blockIdx.x, gridDim.x);
  Lds3D<R> l = carve3d<R>(smem, a.nv, a.nf, blockDim.x, a.ww);
  stage_topology(a, l);
  load_own(a, l, (size_t)ci * a.nv);
  __syncthreads();
  cell_stats(l, a.nv, a.nf, a.st_out, a.vol_out, a.em_out, ci, cell_l0<R>(a.a0, ci));
}

// Winding number of point p about the staged cell (l.nx/ny/nz shifted by sh),
// evaluated by one wave: lanes normalise the directions to the cell's
// vertices once into this wave's LDS table (the same bits a per-face
// normalisation gives, so every face's `denom < 1e-8` decision is the
// reference order's), then split the faces.  The sum is reduced in double.
template <typename R>
__device__ __forceinline__ R winding_wave(const Lds3D<R>& l, Q4<R>* ut, int nv, int nf, V3<R> sh,
                                          V3<R> p) {
  const int lane = threadIdx.x & (DPM_WAVE - 1);
  for (int v = lane; v < nv; v += DPM_WAVE) {
    // Verts + shift - point, in the reference's order (Cell3D_Kernel.cl:285-287)
    const V3<R> A = {(l.nx[v] + sh.x) - p.x, (l.ny[v] + sh.y) - p.y, (l.nz[v] + sh.z) - p.z};
    const V3<R> u = vnormalize(A);
    ut[v] = Q4<R>{u.x, u.y, u.z, R(0)};
  }
  wave_lds_sync();
  R part = R(0);
  for (int f = lane; f < nf; f += DPM_WAVE) {
    int ia, ib, ic;
    face_of(l, f, ia, ib, ic);
    const Q4<R> ua = ut[ia], ub = ut[ib], uc = ut[ic];
    const V3<R> u = {ua.x, ua.y, ua.z};
    const V3<R> v = {ub.x, ub.y, ub.z};
    const V3<R> w = {uc.x, uc.y, uc.z};
    const R denom = ((R(1) + vdot(u, v)) + vdot(v, w)) + vdot(w, u);
    if (denom < R(1e-8f)) continue;
    const R num = vdot(u, vcross(v, w));
    part += R(2) * atan2(num, denom);
  }
  const double tot = wave_sum_d((double)part);  // every lane (DPP tree)
  wave_lds_sync();  // this wave's table reads finish before it is rewritten
  return (R)(tot / (double)(R(4) * R(3.14159265358979323846)));
}

// One overdamped Euler step of every cell (or, without F_MOVE, one force
// evaluation for diagnostics).  blockDim.x >= nv, a multiple of 64.
// NVC, NFC, NTC, WWC: vertex and face counts, threads per workgroup and
// winding waves fixed at compile time (0: taken from the arguments).
template <typename R, int NVC, int NFC, int NTC, int WWC, bool PROBE = true>
__device__ void k3d_step_body(const Step3D<R>& a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int nv = NVC > 0 ? NVC : a.nv, nf = NFC > 0 ? NFC : a.nf, NT = NTC > 0 ? NTC : (int)blockDim.x, tid = threadIdx.x;
  const int ww = WWC > 0 ? WWC : a.ww;
  const int lane = tid & (DPM_WAVE - 1), wave = tid / DPM_WAVE;
  const int ci = xcd_remap(blockIdx.x, gridDim.x);
  Lds3D<R> l = carve3d<R>(smem, nv, nf, NT, ww);

  const int t = a.cell_tissue[ci];
  const int cb = a.tissue_begin[t], ce = a.tissue_begin[t + 1];
  const R L = (R)a.tL[t];
  const int pbc = a.tpbc[t];
  const R Kre = (R)a.tKre[t], Kat = (R)a.tKat[t], ksg = (R)a.tksg[t];
  const CellStat<R> sti = a.st_in[ci];
  const V3<R> Ci = {sti.cx, sti.cy, sti.cz};
  const R Ri = sti.r;
  const R Kvi = (R)a.Kv[ci], Kai = (R)a.Ka[ci], Ksi = (R)a.Ks[ci], Kbi = (R)a.Kb[ci];
  const R v0i = (R)a.v0[ci], a0i = (R)a.a0[ci], l0i = cell_l0<R>(a.a0, ci);
  const unsigned mask = a.mask;
  const size_t base = (size_t)ci * nv;
  const bool own = tid < nv;
  const V3<R> zero = {R(0), R(0), R(0)};

  // per-vertex topology rows in registers
  int vfr[DPM3D_MAXDEG], adjr[DPM3D_MAXDEG];
#pragma unroll
  for (int k = 0; k < DPM3D_MAXDEG; k++) {
    vfr[k] = own ? a.vf[tid * DPM3D_MAXDEG + k] : -1;
    adjr[k] = own ? a.adj[tid * DPM3D_MAXDEG + k] : -1;
  }
  // The candidate scan's statistics (every cell of a tissue of at most NT
  // cells, no cell list) are loaded here with the cell's own data, one
  // memory round trip for both (float instances with compile-time sizes; in
  // the others the registers held until the scan would spill).
  const bool scan_pre = (sizeof(R) == 4 && NTC > 0) &&
                        (((mask & M_CONTACT) && Kre != R(0)) || ((mask & M_ADHESION) && Kat != R(0)));
  const bool scan_all = scan_pre && !(a.bin_list && a.tbin_nb[t] > 0) && ce - cb <= NT;  // block-uniform
  CellStat<R> st_pre{};
  if (scan_all && tid < ce - cb) st_pre = a.st_in[cb + tid];
  if constexpr (NFC > 0 && NVC > 0 && NTC > 0) {
    // compile-time sizes: every load in flight before the first LDS store
    static_assert(NFC <= 2 * NTC, "at most two faces per thread");
    const int f0 = tid, f1 = tid + NTC;
    int a0 = 0, b0 = 0, e0 = 0, a1 = 0, b1 = 0, e1 = 0;
    if (f0 < NFC) {
      a0 = a.faces[3 * f0];
      b0 = a.faces[3 * f0 + 1];
      e0 = a.faces[3 * f0 + 2];
    }
    if (f1 < NFC) {
      a1 = a.faces[3 * f1];
      b1 = a.faces[3 * f1 + 1];
      e1 = a.faces[3 * f1 + 2];
    }
    R px = R(0), py = R(0), pz = R(0);
    if (own) {
      px = a.xin[base + tid];
      py = a.yin[base + tid];
      pz = a.zin[base + tid];
    }
    if (f0 < NFC) l.sf[f0] = make_uint2((unsigned)a0 | ((unsigned)b0 << 16), (unsigned)e0);
    if (f1 < NFC) l.sf[f1] = make_uint2((unsigned)a1 | ((unsigned)b1 << 16), (unsigned)e1);
    if (own) {
      l.sx[tid] = px;
      l.sy[tid] = py;
      l.sz[tid] = pz;
    }
  } else {
    stage_topology(a, l);
    load_own(a, l, base);
  }
  __syncthreads();

  const V3<R> P = own ? lds_v(l.sx, l.sy, l.sz, tid) : zero;
  V3<R> F = zero;
  bool ring_done = false;  // shape forces by the ring walk (bending's first sum in ring_sb)
  V3<R> ring_sb = zero;

  // ---------------------------------------------------------- shape forces
  // Per-face pass: each face's volume, edge-spring and substrate terms for
  // its three corners are computed once into LDS; every vertex then gathers
  // its incident (face, corner) entries in ascending face order.
  if (mask & (M_VOLUME | M_SURFACE | M_SUBSTRATE)) {  // block-uniform
    // Volume strain of the current volume, in double (the residual V/v0 - 1
    // cancels when the cell is near its target).
    const R strain = (R)(a.vol_in[ci] / (double)v0i - 1.0);
    const R kvol = -Kvi * strain / R(6);
    const bool vol_on = (mask & M_VOLUME) && Kvi != R(0);
    const bool surf_on = (mask & M_SURFACE) && !(Kai < R(1e-8f));
    const bool sub_on = (mask & M_SUBSTRATE) && !(Ksi < R(1e-12f));
    const R sc = Kai * sqrt(a0i) / l0i * R(0.3f);
    const R inv_l0 = R(1) / l0i;
    const double sub_inv_l0 = 1.0 / (double)l0i;
    if ((sizeof(R) == 4 || NTC > 0) && a.closed) {  // block-uniform
      // Closed mesh, neighbour rows in ring order (face q is (v, u_q, u_(q+1))):
      // one walk around each vertex's ring, as the wave-per-cell step takes
      // it — volume kvol sum_q d_q x d_(q+1), edge springs 2 sc sum_q s(d_q),
      // bending's first sum, substrate per facing-down incident face.  A
      // ring edge shorter than 2e-12 (the reference skips such faces'
      // springs) sends the cell back to the face pass below.  Float steps
      // and the compile-time-sized instances only: in the general double
      // kernel (128 registers at 1024 threads) the walk's registers push
      // the rest of the step into spills (measured 5 % slower on 4096 x 642
      // cells), and the face pass stays.
      unsigned* dbits = reinterpret_cast<unsigned*>(l.fcx);  // dead until the candidates
      if (sub_on) {
        for (int f0 = wave * DPM_WAVE; f0 < nf; f0 += NT) {  // 64 consecutive faces per wave
          const int f = f0 + lane;
          bool dn = false;
          if (f < nf) {
            int i0, i1, i2;
            face_of(l, f, i0, i1, i2);
            const V3<R> P0 = lds_v(l.sx, l.sy, l.sz, i0);
            dn = facing_down(vcross(vsub(lds_v(l.sx, l.sy, l.sz, i1), P0), vsub(lds_v(l.sx, l.sy, l.sz, i2), P0)));
            if (PROBE && dn && a.p_down) a.p_down[(size_t)ci * nf + f] = 1;
          }
          const unsigned long long m = __ballot(dn);
          if (lane == 0) {
            dbits[f0 / 32] = (unsigned)m;
            if (f0 / 32 + 1 < (nf + 31) / 32) dbits[f0 / 32 + 1] = (unsigned)(m >> 32);
          }
        }
        __syncthreads();
      }
      bool bad = false;
      V3<R> Fr = zero;
      if (own) {
        // (Force arithmetic, its products fused into FMAs; the short-edge
        // test keeps a factor 2 of margin to the reference's 1e-12 cut.)
        const V3<R> d0 = vsub(lds_v(l.sx, l.sy, l.sz, adjr[0]), P);
        V3<R> dq = d0, Fv = zero, Fs = zero, sb = zero;
#pragma unroll
        for (int q = 0; q < DPM3D_MAXDEG; q++) {
          if (surf_on) {
            const R d2 = vdot_f(dq, dq);
            bad = bad || d2 < R(4e-24f);
            Fs = vmadd_f(dq, spring_factor(d2, inv_l0), Fs);
          }
          sb = vadd(sb, dq);
          const int un = q + 1 < DPM3D_MAXDEG ? adjr[q + 1] : -1;
          const V3<R> dn = un < 0 ? d0 : vsub(lds_v(l.sx, l.sy, l.sz, un), P);
          if (vol_on) Fv = vcross_add_f(dq, dn, Fv);
          if (un < 0) break;
          dq = dn;
        }
        Fr = vadd(vscale(Fv, kvol), vscale(Fs, R(2) * sc));
        ring_sb = sb;
        if (sub_on) {
          int ndown = 0;
#pragma unroll
          for (int k = 0; k < DPM3D_MAXDEG; k++) {
            const int e = vfr[k];  // face*3 + corner
            if (e < 0) break;
            const int f = e / 3;
            ndown += (dbits[f >> 5] >> (f & 31)) & 1u;
          }
          if (ndown) Fr = vadd(Fr, vscale(substrate_corner(P, Ci, Ksi, ksg, L, l0i, sub_inv_l0), (R)ndown));
        }
      }
      ring_done = !__syncthreads_or(bad);  // also: every facing-down bit read
      if (ring_done) F = Fr;
    }
    for (int f = tid; f < nf && !ring_done; f += NT) {
      int i0, i1, i2;
      face_of(l, f, i0, i1, i2);
      const V3<R> P0 = lds_v(l.sx, l.sy, l.sz, i0);
      const V3<R> P1 = lds_v(l.sx, l.sy, l.sz, i1);
      const V3<R> P2 = lds_v(l.sx, l.sy, l.sz, i2);
      V3<R> c0 = zero, c1 = zero, c2 = zero;
      if (vol_on) {  // Cell3D_Kernel.cl:95-111
        const V3<R> A = vsub(P1, Ci), B = vsub(P2, Ci), Cc = vsub(P0, Ci);
        c0 = vscale(vcross(A, B), kvol);
        c1 = vscale(vcross(B, Cc), kvol);
        c2 = vscale(vcross(Cc, A), kvol);
      }
      if (surf_on) {  // Cell3D_Kernel.cl:140-176
        const V3<R> lv0 = vsub(P1, P0), lv1 = vsub(P2, P1), lv2 = vsub(P0, P2);
        const R len0 = sqrt(vdot(lv0, lv0)), len1 = sqrt(vdot(lv1, lv1)), len2 = sqrt(vdot(lv2, lv2));
        if (!(len0 < R(1e-12f) || len1 < R(1e-12f) || len2 < R(1e-12f))) {
          const V3<R> e0 = vscale(lv0, (len0 * inv_l0 - R(1)) / len0);
          const V3<R> e1 = vscale(lv1, (len1 * inv_l0 - R(1)) / len1);
          const V3<R> e2 = vscale(lv2, (len2 * inv_l0 - R(1)) / len2);
          c0 = vadd(c0, vscale(vsub(e0, e2), sc));
          c1 = vadd(c1, vscale(vsub(e1, e0), sc));
          c2 = vadd(c2, vscale(vsub(e2, e1), sc));
        }
      }
      if (sub_on) {  // Cell3D_Kernel.cl:206-246
        const V3<R> un = vnormalize(vcross(vsub(P1, P0), vsub(P2, P0)));
        if (un.z < R(-0.1f)) {
          if (PROBE && a.p_down) a.p_down[(size_t)ci * nf + f] = 1;
          c0 = vadd(c0, substrate_corner(P0, Ci, Ksi, ksg, L, l0i, sub_inv_l0));
          c1 = vadd(c1, substrate_corner(P1, Ci, Ksi, ksg, L, l0i, sub_inv_l0));
          c2 = vadd(c2, substrate_corner(P2, Ci, Ksi, ksg, L, l0i, sub_inv_l0));
        }
      }
      l.fcx[3 * f] = c0.x;
      l.fcy[3 * f] = c0.y;
      l.fcz[3 * f] = c0.z;
      l.fcx[3 * f + 1] = c1.x;
      l.fcy[3 * f + 1] = c1.y;
      l.fcz[3 * f + 1] = c1.z;
      l.fcx[3 * f + 2] = c2.x;
      l.fcy[3 * f + 2] = c2.y;
      l.fcz[3 * f + 2] = c2.z;
    }
    if (!ring_done) {  // block-uniform
      __syncthreads();
      if (own) {
#pragma unroll
        for (int k = 0; k < DPM3D_MAXDEG; k++) {
          const int e = vfr[k];  // face*3 + corner
          if (e < 0) break;
          F = vadd(F, V3<R>{l.fcx[e], l.fcy[e], l.fcz[e]});
        }
      }
      __syncthreads();  // the shape-force bytes are reused below
    }
  }

  // --------------------------------------------------------------- bending
  if ((mask & M_BENDING) && Kbi != R(0)) {  // block-uniform
    if (own) {
      V3<R> s = zero;
      if (ring_done) {
        s = ring_sb;
      } else {
#pragma unroll
        for (int k = 0; k < DPM3D_MAXDEG; k++) {
          const int u = adjr[k];
          if (u < 0) break;
          s = vadd(s, vsub(lds_v(l.sx, l.sy, l.sz, u), P));
        }
      }
      l.sb[3 * tid] = s.x;
      l.sb[3 * tid + 1] = s.y;
      l.sb[3 * tid + 2] = s.z;
    }
    __syncthreads();
    if (own) {
      V3<R> acc = zero;
      int deg = 0;
#pragma unroll
      for (int k = 0; k < DPM3D_MAXDEG; k++) {
        const int u = adjr[k];
        if (u < 0) break;
        deg++;
        acc = vadd(acc, V3<R>{l.sb[3 * u], l.sb[3 * u + 1], l.sb[3 * u + 2]});
      }
      const V3<R> sv = {l.sb[3 * tid], l.sb[3 * tid + 1], l.sb[3 * tid + 2]};
      const V3<R> lap = vsub(acc, vscale(sv, (R)deg));
      F = vadd(F, vscale(lap, -Kbi));
    }
  }

  // ------------------------------------------- contact and vertex adhesion
  const bool do_contact = (mask & M_CONTACT) && Kre != R(0);
  const bool do_adh = (mask & M_ADHESION) && Kat != R(0);
  unsigned long long n_wind = 0, n_contact = 0, n_cand = 0;
  R wsum = R(0);  // the contact force is wsum * Kre/2 * normalize(C_i - p)
  int npart = 0;
  if (do_contact || do_adh) {  // block-uniform
    const R L0i = fmax(l0i, R(1e-12f));
    const R invL = R(1) / L;
    Q4<R>* ut = l.utab + (size_t)wave * nv;
    // candidate source: every cell of the tissue in order, or the cells of
    // the 27 buckets around this cell, sorted by index (block-uniform)
    const int* src = nullptr;
    int nsrc = ce - cb;
    if (a.bin_list && a.tbin_nb[t] > 0) {
      const int nb = a.tbin_nb[t], nbuck = a.tbin_nbuck[t];
      const int* bs = a.bin_start + a.tbin_hoff[t];
      const R wb = L / (R)nb;
      int* gid = l.gath;        // [27] bucket ids
      int* gst = l.gath + 32;   // [27] bucket starts
      int* gof = l.gath + 64;   // [28] prefix of bucket sizes
      if (tid < DPM_WAVE) {  // the whole first wave: distinct buckets, their cells' offsets
        int g = 0;
        if (tid < 27) {
          const int bx = bin_coord(Ci.x, L, wb, nb), by = bin_coord(Ci.y, L, wb, nb),
                    bz = bin_coord(Ci.z, L, wb, nb);
          const int dx = tid % 3 - 1, dy = (tid / 3) % 3 - 1, dz = tid / 9 - 1;
          g = (int)bin_bucket((bx + dx + nb) % nb, (by + dy + nb) % nb, (bz + dz + nb) % nb, nb, nbuck);
        }
        const int nu = distinct_buckets27(g, bs, gst, gof);
        if (tid == 0) gid[31] = nu;
      }
      __syncthreads();
      const int nu = gid[31], n = gof[nu];
      if (n <= NT) {  // block-uniform; otherwise scan every cell
        if (tid < n) {
          int u = 0;
          while (gof[u + 1] <= tid) u++;
          l.glist[tid] = a.bin_list[cb + gst[u] + (tid - gof[u])];
        }
        __syncthreads();
        if (tid < n) {  // rank sort (indices are distinct)
          const int v = l.glist[tid];
          int r = 0;
          for (int u = 0; u < n; u++) r += l.glist[u] < v;
          l.clist[r] = v;
        }
        __syncthreads();
        if (tid < n) l.glist[tid] = l.clist[tid];
        __syncthreads();
        src = l.glist;
        nsrc = n;
      }
    }
    for (int c0 = 0; c0 < nsrc; c0 += NT) {
      const int k = c0 + tid;
      const int cj = k < nsrc ? (src ? src[k] : cb + k) : ce;
      bool cand = false;
      CellStat<R> stj{};
      if (cj < ce && cj != ci) {
        if (scan_all && c0 == 0)
          stj = st_pre;
        else
          stj = a.st_in[cj];
        V3<R> sh = zero;
        if (pbc)
          sh = v3(image_shift(Ci.x - stj.cx, L, invL), image_shift(Ci.y - stj.cy, L, invL),
                  image_shift(Ci.z - stj.cz, L, invL));
        const V3<R> d = {stj.cx + sh.x - Ci.x, stj.cy + sh.y - Ci.y, stj.cz + sh.z - Ci.z};
        R reach = R(0);
        if (do_contact) reach = Ri + stj.r + contact_margin(stj.r, stj.em);
        if (do_adh) {
          const R l0m = fmax(L0i, fmax(stj.l0, R(1e-12f)));
          const R r2 = Ri + stj.r + R(2) * l0m;
          reach = fmax(reach, r2);
          // Adhesion takes the minimum image per vertex pair
          // (Cell3D_Kernel.cl:341-346), which can differ from the
          // centre-of-mass image when the box is small: then no culling.
          if (pbc && R(2) * r2 >= R(0.999f) * L) cand = true;
        }
        const R rr = reach * R(1.0001f) + R(1e-6f);
        cand = cand || vdot(d, d) <= rr * rr;
      }
      const int ncand = block_compact_with(cand, cj, l.clist, l.cnt, [&](int slot) {
        if (slot < kCandStats) l.cst[slot] = stj;
      });
      n_cand += ncand;
      // candidate q's statistics (wave-uniform q)
      auto cand_stat = [&](int q) { return q < kCandStats ? l.cst[q] : a.st_in[l.clist[q]]; };
      // Separation along the centre line n = (C_j - C_i) / D (float steps
      // and the compile-time-sized instances): every vertex p of this cell
      // has (p - C_i).n <= hi and every point x of cell j's (shifted) hull
      // has (C_j - x).n <= hj, so |x - p| >= D - hi - hj, less slack for the
      // float sums (1e-4 of the cells' extent, 1e-6 of the coordinates).
      // Contact needs a point closer than the culling margin m_j to a face
      // of j (contact_margin), adhesion a vertex pair closer than 2 l0 — on
      // pairs that take no periodic image (the unshifted centres and radii
      // keep every component below 0.48 L; the shift is 0 then).  The waves
      // take the candidates in turn (maxima by wave shuffles, cell j's
      // vertices read from the L2), so the bound costs no workgroup barrier
      // per candidate; the candidates that need neither pass are dropped
      // and the rest keep their two bits (1 contact, 2 adhesion) above bit
      // 28 of their entry.  In the general double kernel the extra
      // registers cost more than the skipped work (spills, measured on
      // 4096 x 642 cells), and every candidate takes both passes.
      // the candidates that take a pass: entries q | passes << 28 (q the
      // index in l.clist, passes 1 contact, 2 adhesion)
      int nkeep = ncand;
      const int* keep = nullptr;
      bool flagged = false;
      if constexpr (sizeof(R) == 4 || NTC > 0) {
        const int nwaves = NT / DPM_WAVE;
        for (int q = wave; q < ncand; q += nwaves) {  // wave-uniform
          const int cjq = l.clist[q];
          const CellStat<R> stj = cand_stat(q);
          V3<R> sh = zero;
          if (pbc)
            sh = v3(image_shift(Ci.x - stj.cx, L, invL), image_shift(Ci.y - stj.cy, L, invL),
                    image_shift(Ci.z - stj.cz, L, invL));
          const V3<R> Cj = {stj.cx + sh.x, stj.cy + sh.y, stj.cz + sh.z};
          const R l0m = fmax(L0i, fmax(stj.l0, R(1e-12f)));
          const bool all_images = pbc && R(2) * (Ri + stj.r + R(2) * l0m) >= R(0.999f) * L;
          const bool no_image = !pbc || (fabs(Ci.x - stj.cx) + (Ri + stj.r) < R(0.48f) * L &&
                                         fabs(Ci.y - stj.cy) + (Ri + stj.r) < R(0.48f) * L &&
                                         fabs(Ci.z - stj.cz) + (Ri + stj.r) < R(0.48f) * L);
          const V3<R> dC = vsub(Cj, Ci);
          const R D = sqrt(vdot(dC, dC));
          const V3<R> nC = D > R(0) ? vscale(dC, R(1) / D) : zero;
          R hi = R(-3.0e38f), hj = R(-3.0e38f);
          const size_t bj = (size_t)cjq * nv;
          for (int v = lane; v < nv; v += DPM_WAVE) {
            hi = fmax(hi, vdot(vsub(lds_v(l.sx, l.sy, l.sz, v), Ci), nC));
            hj = fmax(hj, vdot(v3(stj.cx - a.xin[bj + v], stj.cy - a.yin[bj + v], stj.cz - a.zin[bj + v]), nC));
          }
          hi = wave_max_all(hi);
          hj = wave_max_all(hj);
          bool skip_contact = false, skip_adh = false;
          const R Mc = ((fabs(Ci.x) + fabs(Ci.y)) + fabs(Ci.z)) + ((fabs(Cj.x) + fabs(Cj.y)) + fabs(Cj.z));
          if (D > R(0)) {
            const R gap = ((D - hi) - hj) - (R(1e-4f) * ((Ri + stj.r) + D) + R(1e-6f) * Mc);
            skip_contact = gap >= contact_margin(stj.r, stj.em);
            skip_adh = !all_images && no_image && gap >= R(2) * l0m;
          }
          const int bits = ((do_contact && !skip_contact) ? 1 : 0) | ((do_adh && !skip_adh) ? 2 : 0);
          if (lane == 0) l.plist[q] = q | (bits << 28);
        }
        __syncthreads();
        const int ent = tid < ncand ? l.plist[tid] : 0;
        nkeep = block_compact(tid < ncand && (ent >> 28) != 0, ent, l.glist, l.cnt);
        keep = l.glist;
        flagged = true;
      }
      // each thread stages one vertex (NT >= nv) of a kept candidate; the
      // next candidate's vertex is loaded while the current one is processed
      auto entry = [&](int qq) { return flagged ? keep[qq] : (qq | (3 << 28)); };
      V3<R> pre = zero;
      if (nkeep > 0 && own) {
        const size_t b0 = (size_t)l.clist[entry(0) & 0x0fffffff] * nv + tid;
        pre = v3(a.xin[b0], a.yin[b0], a.zin[b0]);
      }
      for (int qq = 0; qq < nkeep; qq++) {
        const int ent = entry(qq);
        const int q = ent & 0x0fffffff;
        const int cj2 = l.clist[q];
        const int passes = ent >> 28;
        const CellStat<R> stj = cand_stat(q);
        V3<R> sh = zero;
        if (pbc)
          sh = v3(image_shift(Ci.x - stj.cx, L, invL), image_shift(Ci.y - stj.cy, L, invL),
                  image_shift(Ci.z - stj.cz, L, invL));
        if (own) {
          l.nx[tid] = pre.x;
          l.ny[tid] = pre.y;
          l.nz[tid] = pre.z;
        }
        __syncthreads();
        if (qq + 1 < nkeep && own) {
          const size_t b1 = (size_t)l.clist[entry(qq + 1) & 0x0fffffff] * nv + tid;
          pre = v3(a.xin[b1], a.yin[b1], a.zin[b1]);
        }
        const bool skip_contact = !(passes & 1), skip_adh = !(passes & 2);
        if (do_contact && !skip_contact) {  // block-uniform
          const V3<R> Cj = {stj.cx + sh.x, stj.cy + sh.y, stj.cz + sh.z};
          const R mj = contact_margin(stj.r, stj.em);
          const R rj = (stj.r + mj) * R(1.0001f) + R(1e-6f);
          bool need = false;
          if (own) {
            const V3<R> d = vsub(P, Cj);
            const R d2 = vdot(d, d);
            need = d2 <= rj * rj;
          }
          const int np = block_compact(need, tid, l.plist, l.cnt);
          if (wave < ww) {
            for (int pi = wave; pi < np; pi += ww) {  // wave-uniform
              const int v = l.plist[pi];
              const V3<R> Pv = lds_v(l.sx, l.sy, l.sz, v);
              // a point beyond the margin from j's hull has winding number 0
              R wv = R(0);
              if (!outside_by_margin(l.nx, l.ny, l.nz, nv, stj.cx, stj.cy, stj.cz, Cj, vsub(Pv, Cj), mj, stj.r)) {
                wv = winding_wave(l, ut, nv, nf, sh, Pv);
                if (lane == 0) n_wind++;
              }
              if (lane == 0) l.wb[v] = wv;
            }
          }
          __syncthreads();
          if (need) {
            const R wv = l.wb[tid];
            const size_t pidx = (base + tid) * (size_t)a.ncmax + (size_t)(cj2 - cb);
            if (PROBE && a.p_winding) a.p_winding[pidx] = (float)wv;
            if (!(fabs(wv) < R(1e-6f))) {
              const V3<R> dir = vnormalize(vsub(Ci, P));
              const R mag = fabs(wv) * R(0.5f) * Kre;
              F = vadd(F, vscale(dir, mag));
              n_contact++;
              wsum += fabs(wv);
              if (a.rec_cell) {
                if (npart < DPM3D_MAXPART) {
                  const size_t r = (base + tid) * DPM3D_MAXPART + npart;
                  a.rec_cell[r] = cj2;
                  a.rec_w[r] = wv;
                  a.rec_shift[3 * r] = sh.x;
                  a.rec_shift[3 * r + 1] = sh.y;
                  a.rec_shift[3 * r + 2] = sh.z;
                } else {
                  atomicOr(a.rec_overflow, 1);
                }
                npart++;
              }
              if (PROBE && a.p_contact) a.p_contact[pidx] = 1;
              if (PROBE && a.p_contact_key) {  // count, then key sum of the partner cells
                atomicAdd(&a.p_contact_key[2 * (base + tid)], 1u);
                atomicAdd(&a.p_contact_key[2 * (base + tid) + 1], cell_key(cj2 - cb));
              }
            }
          }
        }
        const R L0j = fmax(stj.l0, R(1e-12f));
        const R l0m = fmax(L0i, L0j);
        const bool all_images = pbc && R(2) * (Ri + stj.r + R(2) * l0m) >= R(0.999f) * L;
        if (do_adh && !skip_adh) {  // block-uniform
          // Vertices of cell j within reach of this cell (a pair closer than
          // 2 l0 has |q - Ci| <= |q - p| + |p - Ci| < 2 l0 + Ri) and vertices
          // of this cell within reach of cell j, the same bound reversed.
          bool near = false, mine = false;
          if (own) {
            const V3<R> ds = {l.nx[tid] + sh.x - Ci.x, l.ny[tid] + sh.y - Ci.y, l.nz[tid] + sh.z - Ci.z};
            const R rk = (Ri + R(2) * l0m) * R(1.0001f) + R(1e-6f);
            near = all_images || vdot(ds, ds) <= rk * rk;
            const V3<R> dp = {P.x - stj.cx - sh.x, P.y - stj.cy - sh.y, P.z - stj.cz - sh.z};
            const R r2 = (stj.r + R(2) * l0m) * R(1.0001f) + R(1e-6f);
            mine = all_images || vdot(dp, dp) <= r2 * r2;
          }
          int nnear = 0, nmine = 0, slot = 0;
          int* jl = l.plist;
          int* il = l.plist + NT;
          block_compact2(near, tid, jl, mine, tid, il, l.cnt, nnear, nmine, slot);
          if (nnear > 0 && nmine > 0) {  // block-uniform
            // (own vertex, slice of near vertices) per thread; each thread
            // sums its slice in order into LDS, then every own vertex adds
            // its slices in order: deterministic, every lane busy.
            const int S = max(1, min(nnear, NT / nmine));
            const int per = (nnear + S - 1) / S;
            const int iv = tid % nmine, sl = tid / nmine;
            Q4<R>* acc = l.utab;
            if (sl < S) {
              const V3<R> Pv = lds_v(l.sx, l.sy, l.sz, il[iv]);
              V3<R> Fa = zero;
              int cnt_own = 0;
              const R half_l = R(0.49f) * L;
              const AdhCut<R> cut_own(L0i * R(2)), cut_rev(L0j * R(2));
              const double kh = 0.5 * (double)Kat, inv_l0i = sti.inv_l0, inv_l0j = stj.inv_l0;
              const float kh_f = (float)kh, inv_l0i_f = (float)inv_l0i, inv_l0j_f = (float)inv_l0j;
              const int m1 = min(nnear, (sl + 1) * per);
              for (int m = sl * per; m < m1; m++) {
                const int vj = jl[m];
                const V3<R> pj = lds_v(l.nx, l.ny, l.nz, vj);
                // the own vertex's loop: delta = pj - p, minimum image per
                // component (Cell3D_Kernel.cl:341-346); a component below
                // 0.49 L rounds to image 0, which leaves it unchanged.
                V3<R> dl = vsub(pj, Pv);
                V3<R> dr = vsub(Pv, pj);  // the partner's loop: delta = p - pj
                V3<R> nl = zero, nr = zero;  // image indices
                bool imaged = false;
                if (pbc && !(fabs(dl.x) < half_l && fabs(dl.y) < half_l && fabs(dl.z) < half_l)) {
                  nl = v3(round(dl.x / L), round(dl.y / L), round(dl.z / L));
                  nr = v3(round(dr.x / L), round(dr.y / L), round(dr.z / L));
                  dl = vsub(dl, v3(L * nl.x, L * nl.y, L * nl.z));
                  dr = vsub(dr, v3(L * nr.x, L * nr.y, L * nr.z));
                  imaged = true;
                }
                // The decisions are the reference's ones (AdhCut).  The force
                // of an accepted pair across the periodic boundary is
                // evaluated in double on the same image, since there the
                // float difference of two near-L numbers keeps few digits of a
                // short vector; without an image (dr = -dl exactly, the same
                // squared distance) a float step takes it in float
                // (adhesion_term_near).
                const R d2 = vdot(dl, dl);
                if constexpr (sizeof(R) == 4) {
                  if (!imaged) {
                    const bool own_on = cut_own(d2), rev_on = cut_rev(d2);
                    if (own_on || rev_on) {
                      const R y = __builtin_amdgcn_rsqf(d2);
                      if (own_on) {
                        Fa = adhesion_term_near(dl, y, kh_f, inv_l0i_f, Fa);
                        cnt_own++;
                        if (PROBE && a.p_adh_hash)
                          atomicAdd(&a.p_adh_hash[2 * (base + il[iv])], vertex_key(pj.x, pj.y, pj.z));
                      }
                      if (rev_on) {
                        Fa = adhesion_term_near(dl, y, kh_f, inv_l0j_f, Fa);
                        if (PROBE && a.p_adh_hash)
                          atomicAdd(&a.p_adh_hash[2 * (base + il[iv]) + 1], vertex_key(pj.x, pj.y, pj.z));
                      }
                    }
                    continue;
                  }
                }
                if (cut_own(d2)) {
                  Fa = vsub(Fa, adhesion_term(pj, Pv, nl, L, Kat, L0i, kh, inv_l0i));
                  cnt_own++;
                  if (PROBE && a.p_adh_hash) atomicAdd(&a.p_adh_hash[2 * (base + il[iv])], vertex_key(pj.x, pj.y, pj.z));
                }
                const R d2r = vdot(dr, dr);
                if (cut_rev(d2r)) {
                  Fa = vadd(Fa, adhesion_term(Pv, pj, nr, L, Kat, L0j, kh, inv_l0j));
                  if (PROBE && a.p_adh_hash)
                    atomicAdd(&a.p_adh_hash[2 * (base + il[iv]) + 1], vertex_key(pj.x, pj.y, pj.z));
                }
              }
              acc[sl * nmine + iv] = Q4<R>{Fa.x, Fa.y, Fa.z, (R)cnt_own};
            }
            __syncthreads();
            if (mine) {
              int cnt_own = 0;
              for (int s2 = 0; s2 < S; s2++) {
                const Q4<R> q4 = acc[s2 * nmine + slot];
                F = vadd(F, V3<R>{q4.x, q4.y, q4.z});
                cnt_own += (int)q4.w;
              }
              if (PROBE && a.p_adh) a.p_adh[base + tid] += cnt_own;
            }
          }
        }
        __syncthreads();
      }
    }
  }

  // ------------------------------------------------------- Euler and stats
  if (own && a.w_out) a.w_out[base + tid] = wsum;
  if (own && a.rec_cell)
    for (int q = npart; q < DPM3D_MAXPART; q++) a.rec_cell[(base + tid) * DPM3D_MAXPART + q] = -1;
  if (own && (a.flags & F_WRITE_FORCES)) {
    a.fx[base + tid] = F.x;
    a.fy[base + tid] = F.y;
    a.fz[base + tid] = F.z;
  }
  if (a.flags & F_MOVE) {  // block-uniform
    V3<R> Pn = P;
    if (own) {
      if (a.xlo) {
        // x + lo + F dt, the displacement rounded once onto the residual
        R ex, ey, ez;
        Pn = v3(two_sum(P.x, fma(F.x, a.dt, a.xlo[base + tid]), ex),
                two_sum(P.y, fma(F.y, a.dt, a.ylo[base + tid]), ey),
                two_sum(P.z, fma(F.z, a.dt, a.zlo[base + tid]), ez));
        a.xlo[base + tid] = ex;
        a.ylo[base + tid] = ey;
        a.zlo[base + tid] = ez;
      } else {
        Pn = v3(fma(F.x, a.dt, P.x), fma(F.y, a.dt, P.y), fma(F.z, a.dt, P.z));
      }
      a.xout[base + tid] = Pn.x;
      a.yout[base + tid] = Pn.y;
      a.zout[base + tid] = Pn.z;
    }
    __syncthreads();  // every read of the step's positions is done
    if (own) {
      l.sx[tid] = Pn.x;
      l.sy[tid] = Pn.y;
      l.sz[tid] = Pn.z;
    }
    cell_stats_own(l, nv, nf, a.st_out, a.vol_out, a.em_out, ci, l0i, own, Pn);
  }
  if (a.counters) {
    // per wave, without a workgroup barrier: the wave's contact count, its
    // winding count (kept by lane 0), and the candidates (every thread holds
    // the workgroup's count; wave 0 adds it).  Spread over
    // DPM_COUNTER_SLOTS slots so that workgroups do not serialise on one
    // address; the host sums the slots.
    unsigned long long c = n_contact;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, DPM_WAVE);
    if (lane == 0) {
      unsigned long long* slot = a.counters + 4 * ((blockIdx.x * (NT / DPM_WAVE) + wave) % DPM_COUNTER_SLOTS);
      if (blockIdx.x == 0 && wave == 0) atomicAdd(&slot[0], 1ull);
      if (n_wind) atomicAdd(&slot[1], n_wind);
      if (c) atomicAdd(&slot[2], c);
      if (wave == 0 && n_cand) atomicAdd(&slot[3], n_cand);
    }
  }
}

// Elastic energy of each cell at the current positions, all in double
// (centre of mass and volume included, so that a float tissue's energy is
// exact for its float positions): the energies whose negative gradients are
// the volume, edge-spring and bending forces above,
//   U = Kv v0/2 (V/v0 - 1)^2 + sum_faces sum_edges sc l0/2 (l/l0 - 1)^2
//       + Kb/2 sum_v |(L x)_v|^2,
// the diagnostic of the precision study (the substrate, contact and
// adhesion forces of the reference are not gradients of an energy).
template <typename R>
__device__ void k3d_energy_body(const Step3D<R>& a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int nv = a.nv, nf = a.nf, NT = blockDim.x, tid = threadIdx.x;
  const int ci = xcd_remap(blockIdx.x, gridDim.x);
  Lds3D<R> l = carve3d<R>(smem, nv, nf, NT, a.ww);
  stage_topology(a, l);
  load_own(a, l, (size_t)ci * nv);
  __syncthreads();
  const double Kv = a.Kv[ci], Ka = a.Ka[ci], Kb = a.Kb[ci], v0 = a.v0[ci], a0 = a.a0[ci],
               l0 = cell_l0<double>(a.a0, ci);
  const double sc = Ka * sqrt(a0) / l0 * (double)0.3f;
  double cx = 0.0, cy = 0.0, cz = 0.0;
  for (int v = tid; v < nv; v += NT) {
    cx += (double)l.sx[v];
    cy += (double)l.sy[v];
    cz += (double)l.sz[v];
  }
  cx = block_sum(cx, l.red) / nv;
  cy = block_sum(cy, l.red) / nv;
  cz = block_sum(cz, l.red) / nv;
  double e = 0.0, vol = 0.0;
  for (int f = tid; f < nf; f += NT) {
    int i[3];
    face_of(l, f, i[0], i[1], i[2]);
    double p[3][3];
    for (int k = 0; k < 3; k++) {
      p[k][0] = (double)l.sx[i[k]] - cx;
      p[k][1] = (double)l.sy[i[k]] - cy;
      p[k][2] = (double)l.sz[i[k]] - cz;
    }
    vol += (p[0][1] * p[1][2] - p[0][2] * p[1][1]) * p[2][0] +
           (p[0][2] * p[1][0] - p[0][0] * p[1][2]) * p[2][1] +
           (p[0][0] * p[1][1] - p[0][1] * p[1][0]) * p[2][2];
    if ((a.mask & M_SURFACE) && !(Ka < 1e-8f))
      for (int k = 0; k < 3; k++) {
        const double* u = p[k];
        const double* w = p[(k + 1) % 3];
        const double dx = w[0] - u[0], dy = w[1] - u[1], dz = w[2] - u[2];
        const double s = sqrt(dx * dx + dy * dy + dz * dz) / l0 - 1.0;
        e += 0.5 * sc * l0 * s * s;
      }
  }
  if ((a.mask & M_BENDING) && Kb != 0.0)
    for (int v = tid; v < nv; v += NT) {
      double s[3] = {0.0, 0.0, 0.0};
      for (int k = 0; k < DPM3D_MAXDEG; k++) {
        const int u = a.adj[v * DPM3D_MAXDEG + k];
        if (u < 0) break;
        s[0] += (double)l.sx[u] - (double)l.sx[v];
        s[1] += (double)l.sy[u] - (double)l.sy[v];
        s[2] += (double)l.sz[u] - (double)l.sz[v];
      }
      e += 0.5 * Kb * (s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    }
  e = block_sum(e, l.red);
  vol = fabs(block_sum(vol, l.red)) / 6.0;
  if (tid == 0) {
    double ev = 0.0;
    if ((a.mask & M_VOLUME) && Kv != 0.0) {
      const double st = vol / v0 - 1.0;
      ev = 0.5 * Kv * v0 * st * st;
    }
    a.energy[ci] = e + ev;
  }
}

// ------------------------------------------------------------------ adjoint
//
// Vector-Jacobian product of one Euler step x' = x + dt F(x; theta), for the
// discrete adjoint of a trajectory:
//   lambda_n = lambda_{n+1} + dt J_x(x_n)^T lambda_{n+1},
//   dL/dtheta += dt J_theta(x_n)^T lambda_{n+1},
// theta = per-cell (Kv, Ka, Kb, Ks).  Every retained term is cell-local:
//   volume    F = -Kv s grad V, s = V/v0 - 1: J^T l = -Kv/v0 g (g.l)
//             - Kv s grad(g.l), g = grad V, with grad(g.l) per face
//             (l_b x x_c + x_b x l_c)/6 at corner a (cyclic), CoM-relative;
//   surface   each face edge a->b is a spring with F_a = sc d u, F_b = -F_a:
//             J^T l at a = -sc M (l_a - l_b), at b the opposite,
//             M = u u^T / l0 + (d/len)(I - u u^T);
//   bending   F = -Kb L L x is linear: J^T l = -Kb L L l;
//   substrate per facing-down face corner (the facing decision and the
//             height cut-offs are piecewise constant): derivatives of
//             ks(x) = Ks (1 + g x_wrapped / L), of |z| and of
//             normalize(p_xy - C) with respect to p and to the centre of
//             mass C (which moves with every vertex by 1/nv);
//   contact   F = sum_j |w_j| Kre/2 normalize(C - p) over the partners j
//             recorded by the forward step.  The winding number of a closed
//             surface would be piecewise constant, but the reference skips
//             faces with denominator < 1e-8 (those subtending more than a
//             hemisphere, near p), so w = (4 pi)^-1 sum over kept faces and
//             dw = -(4 pi)^-1 sum over skipped faces of dOmega_f, with
//             Omega = 2 atan2(N, D): dOmega = 2 (D dN - N dD)/(N^2 + D^2)
//             through the normalised directions; the skipped faces belong to
//             the partner cell, whose share goes to cx/cy/cz atomically;
// vertex adhesion couples cells and is not part of the adjoint.
template <typename R>
struct Vjp3D {
  const R* x;  // positions x_n, all cells
  const R* y;
  const R* z;
  const R* w;  // contact weight W per vertex at step n (nullptr: no contact)
  const int* rec_cell;  // contact partners at step n (see Step3D)
  const R* rec_w;
  const R* rec_shift;
  R* cx;  // [3 * nverts] contributions to other cells' lambda (atomic)
  R* cy;
  R* cz;
  R* lx;       // adjoint: lambda_{n+1} in, lambda_n out (in place)
  R* ly;
  R* lz;
  double* gtheta;  // [ncells*4] dL/d(Kv, Ka, Kb, Ks) accumulated
  const float* Kv;
  const float* Ka;
  const float* Ks;
  const float* Kb;
  const float* v0;
  const float* a0;
  const int* cell_tissue;
  const float* tL;
  const float* tKre;
  const float* tksg;
  const int* faces;
  const int* vf;
  const int* adj;
  int nv, nf;
  R dt;
  unsigned mask;
};

// LDS of the adjoint kernel (byte offsets): positions, lambda, bending
// scratch [3 nv] each, face table, per (face, corner) contributions, and
// reduction scratch — disjoint, as every region is live at once.
struct LdsVjpLayout {
  size_t x, lam, sb, faces, fc, red, total;
};
__host__ __device__ inline LdsVjpLayout ldsvjp_layout(int nv, int nf, int rsize) {
  LdsVjpLayout L;
  const size_t rs = (size_t)rsize;
  size_t o = 0;
  L.x = o;     o += 3u * rs * (size_t)nv;
  L.lam = o;   o += 3u * rs * (size_t)nv;
  L.sb = o;    o += 3u * rs * (size_t)nv;
  o = (o + 7u) & ~(size_t)7u;
  L.faces = o; o += 8u * (size_t)nf;
  o = (o + 15u) & ~(size_t)15u;
  L.fc = o;    o += 9u * rs * (size_t)nf;
  o = (o + 7u) & ~(size_t)7u;
  L.red = o;   o += 8u * 32u;
  L.total = o;
  return L;
}

template <typename R>
__device__ __forceinline__ V3<R> jnorm(V3<R> v, V3<R> l) {
  // N(v) l = (l - (u.l) u)/|v|, u = v/|v|: the Jacobian of normalize(v)
  const R n = sqrt(vdot(v, v));
  if (n == R(0)) return {R(0), R(0), R(0)};
  const V3<R> u = vscale(v, R(1) / n);
  return vscale(vsub(l, vscale(u, vdot(u, l))), R(1) / n);
}

template <typename R>
__device__ void k3d_vjp_body(const Vjp3D<R>& a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int nv = a.nv, nf = a.nf, NT = blockDim.x, tid = threadIdx.x;
  const int ci = xcd_remap(blockIdx.x, gridDim.x);
  const LdsVjpLayout LL = ldsvjp_layout(nv, nf, (int)sizeof(R));
  Lds3D<R> l{};
  l.sx = reinterpret_cast<R*>(smem + LL.x);
  l.sy = l.sx + nv;
  l.sz = l.sy + nv;
  l.nx = reinterpret_cast<R*>(smem + LL.lam);  // lambda
  l.ny = l.nx + nv;
  l.nz = l.ny + nv;
  l.sb = reinterpret_cast<R*>(smem + LL.sb);
  l.sf = reinterpret_cast<uint2*>(smem + LL.faces);
  l.fcx = reinterpret_cast<R*>(smem + LL.fc);
  l.fcy = l.fcx + 3 * nf;
  l.fcz = l.fcy + 3 * nf;
  l.red = reinterpret_cast<double*>(smem + LL.red);
  const size_t base = (size_t)ci * nv;
  const bool own = tid < nv;
  const V3<R> zero = {R(0), R(0), R(0)};
  for (int f = tid; f < nf; f += NT)
    l.sf[f] = make_uint2((unsigned)a.faces[3 * f] | ((unsigned)a.faces[3 * f + 1] << 16),
                         (unsigned)a.faces[3 * f + 2]);
  for (int v = tid; v < nv; v += NT) {
    l.sx[v] = a.x[base + v];
    l.sy[v] = a.y[base + v];
    l.sz[v] = a.z[base + v];
    l.nx[v] = a.lx[base + v];  // lambda
    l.ny[v] = a.ly[base + v];
    l.nz[v] = a.lz[base + v];
  }
  int vfr[DPM3D_MAXDEG], adjr[DPM3D_MAXDEG];
#pragma unroll
  for (int k = 0; k < DPM3D_MAXDEG; k++) {
    vfr[k] = own ? a.vf[tid * DPM3D_MAXDEG + k] : -1;
    adjr[k] = own ? a.adj[tid * DPM3D_MAXDEG + k] : -1;
  }
  __syncthreads();

  const int t = a.cell_tissue[ci];
  const R L = (R)a.tL[t], Kre = (R)a.tKre[t], ksg = (R)a.tksg[t];
  const R Kvi = (R)a.Kv[ci], Kai = (R)a.Ka[ci], Ksi = (R)a.Ks[ci], Kbi = (R)a.Kb[ci];
  const R v0i = (R)a.v0[ci], a0i = (R)a.a0[ci], l0i = cell_l0<R>(a.a0, ci);
  const unsigned mask = a.mask;

  // centre of mass and volume of x_n, as the forward step's statistics
  double px = 0.0, py = 0.0, pz = 0.0;
  for (int v = tid; v < nv; v += NT) {
    px += (double)l.sx[v];
    py += (double)l.sy[v];
    pz += (double)l.sz[v];
  }
  px = block_sum(px, l.red);
  py = block_sum(py, l.red);
  pz = block_sum(pz, l.red);
  const V3<R> C = {(R)(px / nv), (R)(py / nv), (R)(pz / nv)};
  double vol = 0.0, S1 = 0.0;
  const bool vol_on = (mask & M_VOLUME) && Kvi != R(0);
  for (int f = tid; f < nf; f += NT) {
    int i0, i1, i2;
    face_of(l, f, i0, i1, i2);
    const V3<R> P0 = vsub(lds_v(l.sx, l.sy, l.sz, i0), C);
    const V3<R> P1 = vsub(lds_v(l.sx, l.sy, l.sz, i1), C);
    const V3<R> P2 = vsub(lds_v(l.sx, l.sy, l.sz, i2), C);
    vol += (double)vdot(vcross(P0, P1), P2);
    if (vol_on) {  // S1 = sum_i grad_i V . lambda_i
      S1 += (double)vdot(vcross(P1, P2), lds_v(l.nx, l.ny, l.nz, i0)) +
            (double)vdot(vcross(P2, P0), lds_v(l.nx, l.ny, l.nz, i1)) +
            (double)vdot(vcross(P0, P1), lds_v(l.nx, l.ny, l.nz, i2));
    }
  }
  vol = fabs(block_sum(vol, l.red) / 6.0);
  S1 = block_sum(S1, l.red) / 6.0;
  const R strain = (R)(vol / (double)v0i - 1.0);

  // per face-corner adjoint contributions (into the shape-force bytes)
  const bool surf_on = (mask & M_SURFACE) && !(Kai < R(1e-8f));
  const bool sub_on = (mask & M_SUBSTRATE) && !(Ksi < R(1e-12f));
  const R sc = Kai * sqrt(a0i) / l0i * R(0.3f);
  const R kv1 = -Kvi / v0i * (R)S1;  // coefficient of g_j
  const R kv2 = -Kvi * strain;       // coefficient of grad(g.l)_j
  double gKa = 0.0, gKs = 0.0;
  double Lcx = 0.0, Lcy = 0.0, Lcz = 0.0;  // sum of (dF/dC)^T lambda
  for (int f = tid; f < nf; f += NT) {
    int id[3];
    face_of(l, f, id[0], id[1], id[2]);
    V3<R> P[3], Q[3], lam[3], c[3];
    for (int k = 0; k < 3; k++) {
      P[k] = lds_v(l.sx, l.sy, l.sz, id[k]);
      Q[k] = vsub(P[k], C);
      lam[k] = lds_v(l.nx, l.ny, l.nz, id[k]);
      c[k] = zero;
    }
    if (vol_on)
      for (int k = 0; k < 3; k++) {
        const int b = (k + 1) % 3, cc = (k + 2) % 3;
        const V3<R> g = vscale(vcross(Q[b], Q[cc]), R(1) / R(6));
        const V3<R> h = vscale(vadd(vcross(lam[b], Q[cc]), vcross(Q[b], lam[cc])), R(1) / R(6));
        c[k] = vadd(c[k], vadd(vscale(g, kv1), vscale(h, kv2)));
      }
    if (surf_on) {
      const R inv_l0 = R(1) / l0i;
      for (int e = 0; e < 3; e++) {  // edge e from corner e to corner e+1
        const int ea = e, eb = (e + 1) % 3;
        const V3<R> lv = vsub(P[eb], P[ea]);
        const R len = sqrt(vdot(lv, lv));
        if (len < R(1e-12f)) continue;
        const V3<R> u = vscale(lv, R(1) / len);
        const R d = len * inv_l0 - R(1);
        const V3<R> dl = vsub(lam[ea], lam[eb]);
        const R ud = vdot(u, dl);
        // sc M dl, M = u u^T / l0 + (d/len)(I - u u^T)
        const V3<R> Md = vscale(vadd(vscale(u, ud * inv_l0), vscale(vsub(dl, vscale(u, ud)), d / len)), sc);
        c[ea] = vsub(c[ea], Md);
        c[eb] = vadd(c[eb], Md);
        // lambda . F_surf / Ka: the edge pulls a by sc d u and b by -sc d u
        gKa += (double)(sc * d * ud) / (double)Kai;
      }
    }
    if (sub_on) {
      const V3<R> un = vnormalize(vcross(vsub(P[1], P[0]), vsub(P[2], P[0])));
      if (un.z < R(-0.1f)) {
        for (int k = 0; k < 3; k++) {
          const V3<R> pos = P[k], lk = lam[k];
          R ks = Ksi, dks = R(0);  // ks and d ks / d x
          if (ksg != R(0)) {
            const R xw = pos.x - L * floor(pos.x / L);
            ks = Ksi * (R(1) + ksg * xw / L);
            dks = Ksi * ksg / L;
          }
          V3<R> mu = zero;
          R fdotl = R(0);  // lambda . F_corner
          if (pos.z < R(0)) {  // f1 = (0, 0, ks |z|)
            mu.z += -ks * lk.z;
            mu.x += fabs(pos.z) * dks * lk.z;
            fdotl += ks * fabs(pos.z) * lk.z;
          }
          const R height = fabs(pos.z);
          if (height < l0i * R(2)) {  // f2 = normalize(c2v) m
            const V3<R> c2v = {pos.x - C.x, pos.y - C.y, R(0) - C.z};
            const R m = ks * (R(1) - height / l0i);
            const V3<R> n = vnormalize(c2v);
            const V3<R> Nl = jnorm(c2v, lk);
            const R nl = vdot(n, lk);
            mu.x += m * Nl.x + nl * dks * (R(1) - height / l0i);
            mu.y += m * Nl.y;
            mu.z += nl * (-ks * (pos.z < R(0) ? R(-1) : (pos.z > R(0) ? R(1) : R(0))) / l0i);
            Lcx -= (double)(m * Nl.x);
            Lcy -= (double)(m * Nl.y);
            Lcz -= (double)(m * Nl.z);
            fdotl += m * nl;
          }
          c[k] = vadd(c[k], mu);
          gKs += (double)fdotl / (double)Ksi;
        }
      }
    }
    for (int k = 0; k < 3; k++) {
      l.fcx[3 * f + k] = c[k].x;
      l.fcy[3 * f + k] = c[k].y;
      l.fcz[3 * f + k] = c[k].z;
    }
  }
  __syncthreads();
  V3<R> mu = zero;
  if (own) {
#pragma unroll
    for (int k = 0; k < DPM3D_MAXDEG; k++) {
      const int e = vfr[k];
      if (e < 0) break;
      mu = vadd(mu, V3<R>{l.fcx[e], l.fcy[e], l.fcz[e]});
    }
  }

  // bending: mu += -Kb L (L lambda); dL/dKb = -(L lambda).(L x)
  double gKb = 0.0;
  if ((mask & M_BENDING) && Kbi != R(0)) {
    if (own) {
      V3<R> sl = zero, sx = zero;
      const V3<R> P = lds_v(l.sx, l.sy, l.sz, tid), lt = lds_v(l.nx, l.ny, l.nz, tid);
#pragma unroll
      for (int k = 0; k < DPM3D_MAXDEG; k++) {
        const int u = adjr[k];
        if (u < 0) break;
        sl = vadd(sl, vsub(lds_v(l.nx, l.ny, l.nz, u), lt));
        sx = vadd(sx, vsub(lds_v(l.sx, l.sy, l.sz, u), P));
      }
      l.sb[3 * tid] = sl.x;
      l.sb[3 * tid + 1] = sl.y;
      l.sb[3 * tid + 2] = sl.z;
      gKb = -(double)vdot(sl, sx);
    }
    __syncthreads();
    if (own) {
      V3<R> acc = zero;
      int deg = 0;
#pragma unroll
      for (int k = 0; k < DPM3D_MAXDEG; k++) {
        const int u = adjr[k];
        if (u < 0) break;
        deg++;
        acc = vadd(acc, V3<R>{l.sb[3 * u], l.sb[3 * u + 1], l.sb[3 * u + 2]});
      }
      const V3<R> sv = {l.sb[3 * tid], l.sb[3 * tid + 1], l.sb[3 * tid + 2]};
      mu = vadd(mu, vscale(vsub(acc, vscale(sv, (R)deg)), -Kbi));
    }
  }

  // contact: F = W Kre/2 normalize(C - p), W = sum_j |w_j(p, x_j)|
  if (own && a.w && (mask & M_CONTACT)) {
    const R W = a.w[base + tid];
    if (W != R(0)) {
      const V3<R> P = lds_v(l.sx, l.sy, l.sz, tid), lt = lds_v(l.nx, l.ny, l.nz, tid);
      const V3<R> Nl = jnorm(vsub(C, P), lt);
      const R k = W * R(0.5f) * Kre;
      mu = vsub(mu, vscale(Nl, k));
      Lcx += (double)(k * Nl.x);
      Lcy += (double)(k * Nl.y);
      Lcz += (double)(k * Nl.z);
      // the winding numbers' own dependence (skipped faces only)
      const R nl = vdot(vnormalize(vsub(C, P)), lt);
      const R inv4pi = R(1) / (R(4) * R(3.14159265358979323846));
      for (int q = 0; q < DPM3D_MAXPART; q++) {
        const size_t r = (base + tid) * DPM3D_MAXPART + q;
        const int cj = a.rec_cell[r];
        if (cj < 0) break;
        const R wq = a.rec_w[r];
        const R coef = (wq > R(0) ? R(1) : R(-1)) * R(0.5f) * Kre * nl * inv4pi;
        const V3<R> sh = {a.rec_shift[3 * r], a.rec_shift[3 * r + 1], a.rec_shift[3 * r + 2]};
        const size_t bj = (size_t)cj * nv;
        for (int f = 0; f < nf; f++) {
          int ia, ib, ic;
          face_of(l, f, ia, ib, ic);
          // the forward step's operation order (Cell3D_Kernel.cl:285-287)
          const V3<R> A = {(a.x[bj + ia] + sh.x) - P.x, (a.y[bj + ia] + sh.y) - P.y, (a.z[bj + ia] + sh.z) - P.z};
          const V3<R> B = {(a.x[bj + ib] + sh.x) - P.x, (a.y[bj + ib] + sh.y) - P.y, (a.z[bj + ib] + sh.z) - P.z};
          const V3<R> Cc = {(a.x[bj + ic] + sh.x) - P.x, (a.y[bj + ic] + sh.y) - P.y, (a.z[bj + ic] + sh.z) - P.z};
          const V3<R> u = vnormalize(A), v = vnormalize(B), w = vnormalize(Cc);
          const R den = ((R(1) + vdot(u, v)) + vdot(v, w)) + vdot(w, u);
          if (!(den < R(1e-8f))) continue;  // kept face: no contribution
          const R num = vdot(u, vcross(v, w));
          const R s2 = R(2) / (num * num + den * den);
          const R cN = s2 * den, cD = -s2 * num;
          const V3<R> du = vadd(vscale(vcross(v, w), cN), vscale(vadd(v, w), cD));
          const V3<R> dv = vadd(vscale(vcross(w, u), cN), vscale(vadd(u, w), cD));
          const V3<R> dw = vadd(vscale(vcross(u, v), cN), vscale(vadd(u, v), cD));
          const V3<R> ga = jnorm(A, du), gb = jnorm(B, dv), gc = jnorm(Cc, dw);
          // dw/dp = +(ga + gb + gc)/(4 pi), dw/dx_a = -ga/(4 pi) (skipped face)
          mu = vadd(mu, vscale(vadd(vadd(ga, gb), gc), coef));
          atomicAdd(&a.cx[bj + ia], -coef * ga.x);
          atomicAdd(&a.cy[bj + ia], -coef * ga.y);
          atomicAdd(&a.cz[bj + ia], -coef * ga.z);
          atomicAdd(&a.cx[bj + ib], -coef * gb.x);
          atomicAdd(&a.cy[bj + ib], -coef * gb.y);
          atomicAdd(&a.cz[bj + ib], -coef * gb.z);
          atomicAdd(&a.cx[bj + ic], -coef * gc.x);
          atomicAdd(&a.cy[bj + ic], -coef * gc.y);
          atomicAdd(&a.cz[bj + ic], -coef * gc.z);
        }
      }
    }
  }
  Lcx = block_sum(Lcx, l.red);
  Lcy = block_sum(Lcy, l.red);
  Lcz = block_sum(Lcz, l.red);
  gKa = block_sum(gKa, l.red);
  gKs = block_sum(gKs, l.red);
  gKb = block_sum(gKb, l.red);
  if (own) {
    // the centre of mass moves with each vertex by 1/nv
    mu = vadd(mu, V3<R>{(R)(Lcx / nv), (R)(Lcy / nv), (R)(Lcz / nv)});
    const V3<R> lt = lds_v(l.nx, l.ny, l.nz, tid);
    a.lx[base + tid] = lt.x + a.dt * mu.x;
    a.ly[base + tid] = lt.y + a.dt * mu.y;
    a.lz[base + tid] = lt.z + a.dt * mu.z;
  }
  if (tid == 0) {
    const double dt = (double)a.dt;
    a.gtheta[4 * ci + 0] += dt * (vol_on ? -(double)strain * S1 : 0.0);
    a.gtheta[4 * ci + 1] += dt * gKa;
    a.gtheta[4 * ci + 2] += dt * gKb;
    a.gtheta[4 * ci + 3] += dt * gKs;
  }
}

extern "C" __global__ void __launch_bounds__(1024) k3d_stats_f32(const Step3D<float> a) { k3d_stats_body(a); }
extern "C" __global__ void __launch_bounds__(1024) k3d_stats_f64(const Step3D<double> a) { k3d_stats_body(a); }
extern "C" __global__ void __launch_bounds__(1024) k3d_step_f32(const Step3D<float> a) { k3d_step_body<float, 0, 0, 0, 0>(a); }
extern "C" __global__ void __launch_bounds__(1024) k3d_step_f64(const Step3D<double> a) { k3d_step_body<double, 0, 0, 0, 0>(a); }
// Instances for the icospheres of subdivision 2 and 3 (162 / 642 vertices):
// sizes, workgroup size and winding waves fixed at compile time, so LDS
// offsets are immediates and the register budget is that of 256 or 704
// threads instead of 1024 (the general kernel spills at 128 registers).
// The host takes them where the handle's nv, nf, NT and ww match.
extern "C" __global__ void __launch_bounds__(256) k3d_step_f32_v162(const Step3D<float> a) {
  k3d_step_body<float, 162, 320, 256, 4, false>(a);
}
// the same with the diagnostic outputs (dpm3d_probe*) compiled in
extern "C" __global__ void __launch_bounds__(256) k3d_step_f32_v162_probe(const Step3D<float> a) {
  k3d_step_body<float, 162, 320, 256, 4, true>(a);
}
extern "C" __global__ void __launch_bounds__(704) k3d_step_f32_v642(const Step3D<float> a) { k3d_step_body<float, 642, 1280, 704, 4>(a); }
extern "C" __global__ void __launch_bounds__(704) k3d_step_f64_v642(const Step3D<double> a) { k3d_step_body<double, 642, 1280, 704, 2>(a); }
extern "C" __global__ void __launch_bounds__(256) k3d_step_f64_v162(const Step3D<double> a) { k3d_step_body<double, 162, 320, 256, 4>(a); }
extern "C" __global__ void __launch_bounds__(1024) k3d_energy_f32(const Step3D<float> a) { k3d_energy_body(a); }
extern "C" __global__ void __launch_bounds__(1024) k3d_energy_f64(const Step3D<double> a) { k3d_energy_body(a); }
// lambda += dt * (contributions from other cells), contributions cleared.
template <typename R>
__device__ void k3d_vjp_cross_body(R* lx, R* ly, R* lz, R* cx, R* cy, R* cz, R dt, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    lx[i] += dt * cx[i];
    ly[i] += dt * cy[i];
    lz[i] += dt * cz[i];
    cx[i] = R(0);
    cy[i] = R(0);
    cz[i] = R(0);
  }
}
extern "C" __global__ void k3d_vjp_cross_f32(float* lx, float* ly, float* lz, float* cx, float* cy, float* cz,
                                             float dt, size_t n) {
  k3d_vjp_cross_body(lx, ly, lz, cx, cy, cz, dt, n);
}
extern "C" __global__ void k3d_vjp_cross_f64(double* lx, double* ly, double* lz, double* cx, double* cy,
                                             double* cz, double dt, size_t n) {
  k3d_vjp_cross_body(lx, ly, lz, cx, cy, cz, dt, n);
}
extern "C" __global__ void __launch_bounds__(1024) k3d_vjp_f32(const Vjp3D<float> a) { k3d_vjp_body(a); }
extern "C" __global__ void __launch_bounds__(1024) k3d_vjp_f64(const Vjp3D<double> a) { k3d_vjp_body(a); }
// the double adjoint for workgroups of 256 threads (meshes of up to 256
// vertices): the register budget of 256 threads, where the general kernel
// spills at 128 registers
extern "C" __global__ void __launch_bounds__(256) k3d_vjp_f64_t256(const Vjp3D<double> a) { k3d_vjp_body(a); }
extern "C" __global__ void __launch_bounds__(1024) k3d_bin_f32(const Bin3D<float> a) { k3d_bin_body(a); }
extern "C" __global__ void __launch_bounds__(1024) k3d_bin_f64(const Bin3D<double> a) { k3d_bin_body(a); }
