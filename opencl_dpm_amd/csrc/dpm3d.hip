// 3D DPM timestep kernels for gfx950.
//
// One workgroup owns one cell for a whole step (grid = cells of all tissues,
// workgroups spread over the 8 XCDs so that neighbouring cells share an L2).
// The cell's vertices and the mesh's face table live in LDS for the step.
// Every force term of the reference's per-face kernels is evaluated by
// *gathering* over the faces incident to a vertex in a fixed order, so no
// float atomics are used and a step is bit-reproducible run to run (the
// reference scatters with a compare-and-swap float add,
// Cell3D_Kernel.cl:10-32, whose order varies).
//
// Positions are double-buffered (read set -> write set) and the per-cell
// statistics the next step needs (centre of mass, bounding radius, volume)
// are reduced from the new positions in the same launch, so a step is ONE
// kernel launch reading and writing each vertex once.
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
#define DPM3D_MAXCLUS 64

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

struct Step3D {
  const float* xin;
  const float* yin;
  const float* zin;
  float* xout;
  float* yout;
  float* zout;
  float* fx;
  float* fy;
  float* fz;
  const float4* st_in;  // per cell: COM xyz, bounding radius
  const double* vol_in;
  const float* em_in;   // per cell: longest edge
  float4* st_out;
  double* vol_out;
  float* em_out;
  const float* Kv;
  const float* Ka;
  const float* Ks;
  const float* Kb;
  const float* v0;
  const float* a0;
  const float* l0;
  const int* cell_tissue;
  const int* tissue_begin;
  const float* tL;
  const int* tpbc;
  const float* tKre;
  const float* tKat;
  const float* tksg;
  const int* faces;        // [nf*3]
  const int* vf;           // [nv*MAXDEG] incident (face*3 + corner), ascending, -1 pad
  const int* adj;          // [nv*MAXDEG] neighbour vertices, ascending, -1 pad
  const int* clus_of;      // [nv] cluster of each vertex (cluster k is seeded at vertex k)
  const int* clus_begin;   // [nclus+1]
  const int* clus_members; // [nv] vertices grouped by cluster
  int nclus;
  int nv, nf, ncells;
  int ww;                  // waves that evaluate winding numbers (LDS tables)
  float dt;
  unsigned mask;
  int flags;
  unsigned long long* counters;  // [4]: steps, windings, contacts, candidates
  // diagnostics (nullptr when not probing)
  unsigned char* p_contact;
  float* p_winding;
  unsigned char* p_down;
  int* p_adh;
  int ncmax;
};

// Culling reach added to a cell's bounding radius for the winding-number
// test.  Every face of cell j lies inside its bounding sphere R_j, so a
// point outside R_j + m_j is at least m_j from every face.  A triangle whose
// edges are all <= e fits in a disc of radius e/sqrt(3); seen from distance
// >= e it subtends at most 2*pi*(1 - sqrt(3/4)) ~ 0.84 sr, far from the
// more-than-hemisphere (> pi) solid angle the reference skips
// (denom < 1e-8).  With m_j = e_max(j) no face is skipped, the exact winding
// number of such a point is 0, and its float evaluation (|Omega_f| small,
// about 1e-7 relative rounding each) stays orders of magnitude below the
// 1e-6 pair threshold — the pair is absent in the reference arithmetic too.
__device__ __forceinline__ float contact_margin(float Rj, float emax_j) {
  return emax_j * 1.001f + 1e-6f * Rj;
}

// LDS layout (byte offsets), shared by host (sizing) and device (carving).
struct Lds3DLayout {
  size_t own, nbr, bend, wnum, faces, members, clist, plist, cnt, crad, utab, fc, red, total;
};
__host__ __device__ inline Lds3DLayout lds3d_layout(int nv, int nf, int nt, int ww) {
  Lds3DLayout L;
  size_t o = 0;
  L.own = o;     o += 12u * (size_t)nv;
  L.nbr = o;     o += 12u * (size_t)nv;
  L.bend = o;    o += 12u * (size_t)nv;
  L.wnum = o;    o += 4u * (size_t)nv;
  L.faces = o;   o += 8u * (size_t)nf;                 // (i0 | i1 << 16, i2) per face
  L.members = o; o += 2u * (size_t)nv;
  o = (o + 3u) & ~(size_t)3u;
  L.clist = o;   o += 4u * (size_t)nt;
  L.plist = o;   o += 4u * (size_t)nt;
  L.cnt = o;     o += 4u * 32u;
  L.crad = o;    o += 4u * (3u * DPM3D_MAXCLUS + 1u);  // radii (uint bits), member offsets, active list
  o = (o + 15u) & ~(size_t)15u;
  L.utab = o;    o += 16u * (size_t)nv * (size_t)ww;
  L.fc = o;      o += 36u * (size_t)nf;                // per (face, corner) shape forces
  o = (o + 7u) & ~(size_t)7u;
  L.red = o;     o += 8u * 32u;
  L.total = o;
  return L;
}

struct Lds3D {
  float *sx, *sy, *sz, *nx, *ny, *nz, *sb, *wb, *fcx, *fcy, *fcz;
  float4* utab;
  uint2* sf;
  unsigned short* smem_members;
  int *clist, *plist, *cnt, *cbeg, *cact;
  unsigned* crad;
  double* red;
};
__device__ __forceinline__ Lds3D carve3d(unsigned char* smem, int nv, int nf, int nt, int ww) {
  const Lds3DLayout L = lds3d_layout(nv, nf, nt, ww);
  Lds3D l;
  float* o = reinterpret_cast<float*>(smem + L.own);
  l.sx = o;
  l.sy = o + nv;
  l.sz = o + 2 * nv;
  float* n = reinterpret_cast<float*>(smem + L.nbr);
  l.nx = n;
  l.ny = n + nv;
  l.nz = n + 2 * nv;
  l.sb = reinterpret_cast<float*>(smem + L.bend);
  l.wb = reinterpret_cast<float*>(smem + L.wnum);
  l.sf = reinterpret_cast<uint2*>(smem + L.faces);
  l.smem_members = reinterpret_cast<unsigned short*>(smem + L.members);
  l.clist = reinterpret_cast<int*>(smem + L.clist);
  l.plist = reinterpret_cast<int*>(smem + L.plist);
  l.cnt = reinterpret_cast<int*>(smem + L.cnt);
  l.crad = reinterpret_cast<unsigned*>(smem + L.crad);
  l.cbeg = reinterpret_cast<int*>(smem + L.crad) + DPM3D_MAXCLUS;
  l.cact = l.cbeg + DPM3D_MAXCLUS + 1;
  l.utab = reinterpret_cast<float4*>(smem + L.utab);
  l.fcx = reinterpret_cast<float*>(smem + L.fc);
  l.fcy = l.fcx + 3 * nf;
  l.fcz = l.fcy + 3 * nf;
  l.red = reinterpret_cast<double*>(smem + L.red);
  return l;
}

// Orders this wave's LDS writes before its later LDS reads (the LDS unit
// executes one wave's operations in order; this keeps the compiler from
// reordering them and waits for outstanding LDS traffic).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ f3 lds_v(const float* x, const float* y, const float* z, int i) {
  return mk(x[i], y[i], z[i]);
}

__device__ __forceinline__ void face_of(const Lds3D& l, int f, int& i0, int& i1, int& i2) {
  const uint2 w = l.sf[f];
  i0 = (int)(w.x & 0xffffu);
  i1 = (int)(w.x >> 16);
  i2 = (int)w.y;
}

// Centre of mass (double sums), bounding radius and volume of the cell whose
// vertices are in l.sx/sy/sz.  The volume is summed over faces relative to
// the centre of mass (translation-invariant for a closed surface), per-face
// terms in float, the sum in double — the reference sums absolute-coordinate
// triple products in float (getVolume, Cell3D_Kernel.cl:46-64), which
// cancels catastrophically far from the origin.
__device__ void cell_stats(const Lds3D& l, int nv, int nf, float4* st_out, double* vol_out,
                           float* em_out, int ci) {
  const int tid = threadIdx.x, NT = blockDim.x;
  double px = 0.0, py = 0.0, pz = 0.0;
  for (int v = tid; v < nv; v += NT) {
    px += (double)l.sx[v];
    py += (double)l.sy[v];
    pz += (double)l.sz[v];
  }
  px = block_sum(px, l.red);
  py = block_sum(py, l.red);
  pz = block_sum(pz, l.red);
  const f3 C = {(float)(px / nv), (float)(py / nv), (float)(pz / nv)};
  float r = 0.0f;
  for (int v = tid; v < nv; v += NT) {
    const f3 d = sub(lds_v(l.sx, l.sy, l.sz, v), C);
    r = fmaxf(r, sqrtf(dot(d, d)));
  }
  r = block_max(r, reinterpret_cast<float*>(l.red));
  double vol = 0.0;
  float e2 = 0.0f;
  for (int f = tid; f < nf; f += NT) {
    int i0, i1, i2;
    face_of(l, f, i0, i1, i2);
    const f3 P0 = sub(lds_v(l.sx, l.sy, l.sz, i0), C);
    const f3 P1 = sub(lds_v(l.sx, l.sy, l.sz, i1), C);
    const f3 P2 = sub(lds_v(l.sx, l.sy, l.sz, i2), C);
    vol += (double)dot(cross(P0, P1), P2);
    const f3 e0 = sub(P1, P0), e1 = sub(P2, P1), ee = sub(P0, P2);
    e2 = fmaxf(e2, fmaxf(dot(e0, e0), fmaxf(dot(e1, e1), dot(ee, ee))));
  }
  vol = block_sum(vol, l.red);
  e2 = block_max(e2, reinterpret_cast<float*>(l.red));
  if (tid == 0) {
    // Slack of a few ulp so that the culling tests are never tighter than the
    // largest float distance a vertex can have from C, or the longest edge.
    st_out[ci] = make_float4(C.x, C.y, C.z, r * 1.0000005f + 1e-30f);
    vol_out[ci] = fabs(vol / 6.0);
    em_out[ci] = sqrtf(e2) * 1.000001f;
  }
}

__device__ __forceinline__ void stage_topology(const Step3D& a, const Lds3D& l) {
  const int tid = threadIdx.x, NT = blockDim.x;
  for (int f = tid; f < a.nf; f += NT)
    l.sf[f] = make_uint2((unsigned)a.faces[3 * f] | ((unsigned)a.faces[3 * f + 1] << 16),
                         (unsigned)a.faces[3 * f + 2]);
  for (int i = tid; i < a.nv; i += NT) l.smem_members[i] = (unsigned short)a.clus_members[i];
  if (tid <= a.nclus) l.cbeg[tid] = a.clus_begin[tid];
}

// Kat/2 (d/L0 - 1) delta/d with delta = (b - a) - L n, the adhesion term of
// Cell3D_Kernel.cl:354-357, evaluated in double.
__device__ __forceinline__ f3 adhesion_term(f3 b, f3 a, f3 n, float L, float Kat, float L0) {
  const double Ld = (double)L;
  const double dx = ((double)b.x - (double)a.x) - Ld * (double)n.x;
  const double dy = ((double)b.y - (double)a.y) - Ld * (double)n.y;
  const double dz = ((double)b.z - (double)a.z) - Ld * (double)n.z;
  const double d = sqrt(dx * dx + dy * dy + dz * dz);
  const double k = (double)Kat * 0.5 * (d / (double)L0 - 1.0) / d;
  return mk((float)(k * dx), (float)(k * dy), (float)(k * dz));
}

// Substrate force on one corner of a facing-down face
// (Cell3D_Kernel.cl:229-245): a z < 0 penalty and the spreading push away
// from the centre of mass within 2 l0 of the plane.
__device__ __forceinline__ f3 substrate_corner(f3 pos, f3 C, float Ks, float ksg, float L, float l0) {
  f3 c = {0.0f, 0.0f, 0.0f};
  float ks = Ks;
  if (ksg != 0.0f) {
    const float xw = pos.x - L * floorf(pos.x / L);
    ks = Ks * (1.0f + ksg * xw / L);
  }
  if (pos.z < 0.0f) c.z += ks * fabsf(pos.z);
  const float height = fabsf(pos.z);
  if (height < l0 * 2.0f) {
    const f3 c2v = mk(pos.x - C.x, pos.y - C.y, 0.0f - C.z);
    const float dz = pos.z - 0.0f;
    const float dist = sqrtf((0.0f * 0.0f + 0.0f * 0.0f) + dz * dz);
    const float ftmp = (float)((double)ks * (1.0 - (double)(dist / l0)));
    c = add(c, scale(normalize(c2v), ftmp));
  }
  return c;
}

extern "C" __global__ void __launch_bounds__(1024) k3d_stats(const Step3D a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int nv = a.nv, NT = blockDim.x, tid = threadIdx.x;
  const int ci = xcd_remap(blockIdx.x, gridDim.x);
  Lds3D l = carve3d(smem, nv, a.nf, NT, a.ww);
  const size_t base = (size_t)ci * nv;
  stage_topology(a, l);
  for (int v = tid; v < nv; v += NT) {
    l.sx[v] = a.xin[base + v];
    l.sy[v] = a.yin[base + v];
    l.sz[v] = a.zin[base + v];
  }
  __syncthreads();
  cell_stats(l, nv, a.nf, a.st_out, a.vol_out, a.em_out, ci);
}

// Winding number of point p about the staged cell (l.nx/ny/nz shifted by sh),
// evaluated by one wave: lanes normalise the directions to the cell's
// vertices once into this wave's LDS table (the same bits a per-face
// normalisation gives, so every face's `denom < 1e-8` decision is the
// reference order's), then split the faces.  The sum is reduced in double.
__device__ __forceinline__ float winding_wave(const Lds3D& l, float4* ut,
                                              int nv, int nf, f3 sh, f3 p) {
  const int lane = threadIdx.x & (DPM_WAVE - 1);
  for (int v = lane; v < nv; v += DPM_WAVE) {
    // Verts + shift - point, in the reference's order (Cell3D_Kernel.cl:285-287)
    const f3 A = mk((l.nx[v] + sh.x) - p.x, (l.ny[v] + sh.y) - p.y, (l.nz[v] + sh.z) - p.z);
    const f3 u = normalize(A);
    ut[v] = make_float4(u.x, u.y, u.z, 0.0f);
  }
  wave_lds_sync();
  float part = 0.0f;
  for (int f = lane; f < nf; f += DPM_WAVE) {
    int ia, ib, ic;
    face_of(l, f, ia, ib, ic);
    const float4 ua = ut[ia], ub = ut[ib], uc = ut[ic];
    const f3 u = mk(ua.x, ua.y, ua.z);
    const f3 v = mk(ub.x, ub.y, ub.z);
    const f3 w = mk(uc.x, uc.y, uc.z);
    const float denom = ((1.0f + dot(u, v)) + dot(v, w)) + dot(w, u);
    if (denom < 1e-8f) continue;
    const float num = dot(u, cross(v, w));
    part += 2.0f * atan2f(num, denom);
  }
  const double tot = wave_sum((double)part);
  wave_lds_sync();  // this wave's table reads finish before it is rewritten
  return (float)(tot / (double)(4.0f * 3.14159265358979323846f));
}

// One overdamped Euler step of every cell (or, without F_MOVE, one force
// evaluation for diagnostics).  blockDim.x >= nv, a multiple of 64.
extern "C" __global__ void __launch_bounds__(1024) k3d_step(const Step3D a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int nv = a.nv, nf = a.nf, NT = blockDim.x, tid = threadIdx.x;
  const int lane = tid & (DPM_WAVE - 1), wave = tid / DPM_WAVE;
  const int ci = xcd_remap(blockIdx.x, gridDim.x);
  Lds3D l = carve3d(smem, nv, nf, NT, a.ww);

  const int t = a.cell_tissue[ci];
  const int cb = a.tissue_begin[t], ce = a.tissue_begin[t + 1];
  const float L = a.tL[t];
  const int pbc = a.tpbc[t];
  const float Kre = a.tKre[t], Kat = a.tKat[t], ksg = a.tksg[t];
  const float4 sti = a.st_in[ci];
  const f3 Ci = {sti.x, sti.y, sti.z};
  const float Ri = sti.w;
  const float Kvi = a.Kv[ci], Kai = a.Ka[ci], Ksi = a.Ks[ci], Kbi = a.Kb[ci];
  const float v0i = a.v0[ci], a0i = a.a0[ci], l0i = a.l0[ci];
  const unsigned mask = a.mask;
  const size_t base = (size_t)ci * nv;
  const bool own = tid < nv;

  // per-vertex topology rows in registers
  int vfr[DPM3D_MAXDEG], adjr[DPM3D_MAXDEG];
#pragma unroll
  for (int k = 0; k < DPM3D_MAXDEG; k++) {
    vfr[k] = own ? a.vf[tid * DPM3D_MAXDEG + k] : -1;
    adjr[k] = own ? a.adj[tid * DPM3D_MAXDEG + k] : -1;
  }
  stage_topology(a, l);
  for (int v = tid; v < nv; v += NT) {
    l.sx[v] = a.xin[base + v];
    l.sy[v] = a.yin[base + v];
    l.sz[v] = a.zin[base + v];
  }
  __syncthreads();

  const f3 P = own ? lds_v(l.sx, l.sy, l.sz, tid) : mk(0.0f, 0.0f, 0.0f);
  f3 F = {0.0f, 0.0f, 0.0f};

  // ---------------------------------------------------------- shape forces
  // Per-face pass: each face's volume, edge-spring and substrate terms for
  // its three corners are computed once into LDS; every vertex then gathers
  // its incident (face, corner) entries in ascending face order.
  if (mask & (M_VOLUME | M_SURFACE | M_SUBSTRATE)) {  // block-uniform
    // Volume strain of the current volume, in double (the residual V/v0 - 1
    // cancels when the cell is near its target).
    const float strain = (float)(a.vol_in[ci] / (double)v0i - 1.0);
    const float kvol = -Kvi * strain / 6.0f;
    const bool vol_on = (mask & M_VOLUME) && Kvi != 0.0f;
    const bool surf_on = (mask & M_SURFACE) && !(Kai < 1e-8f);
    const bool sub_on = (mask & M_SUBSTRATE) && !(Ksi < 1e-12f);
    const float sc = Kai * sqrtf(a0i) / l0i * 0.3f;
    const float inv_l0 = 1.0f / l0i;
    for (int f = tid; f < nf; f += NT) {
      int i0, i1, i2;
      face_of(l, f, i0, i1, i2);
      const f3 P0 = lds_v(l.sx, l.sy, l.sz, i0);
      const f3 P1 = lds_v(l.sx, l.sy, l.sz, i1);
      const f3 P2 = lds_v(l.sx, l.sy, l.sz, i2);
      f3 c0 = {0.0f, 0.0f, 0.0f}, c1 = c0, c2 = c0;
      if (vol_on) {  // Cell3D_Kernel.cl:95-111
        const f3 A = sub(P1, Ci), B = sub(P2, Ci), Cc = sub(P0, Ci);
        c0 = scale(cross(A, B), kvol);
        c1 = scale(cross(B, Cc), kvol);
        c2 = scale(cross(Cc, A), kvol);
      }
      if (surf_on) {  // Cell3D_Kernel.cl:140-176
        const f3 lv0 = sub(P1, P0), lv1 = sub(P2, P1), lv2 = sub(P0, P2);
        const float len0 = sqrtf(dot(lv0, lv0)), len1 = sqrtf(dot(lv1, lv1)),
                    len2 = sqrtf(dot(lv2, lv2));
        if (!(len0 < 1e-12f || len1 < 1e-12f || len2 < 1e-12f)) {
          const f3 e0 = scale(lv0, (len0 * inv_l0 - 1.0f) / len0);
          const f3 e1 = scale(lv1, (len1 * inv_l0 - 1.0f) / len1);
          const f3 e2 = scale(lv2, (len2 * inv_l0 - 1.0f) / len2);
          c0 = add(c0, scale(sub(e0, e2), sc));
          c1 = add(c1, scale(sub(e1, e0), sc));
          c2 = add(c2, scale(sub(e2, e1), sc));
        }
      }
      if (sub_on) {  // Cell3D_Kernel.cl:206-246
        const f3 un = normalize(cross(sub(P1, P0), sub(P2, P0)));
        if (un.z < -0.1f) {
          if (a.p_down) a.p_down[(size_t)ci * nf + f] = 1;
          c0 = add(c0, substrate_corner(P0, Ci, Ksi, ksg, L, l0i));
          c1 = add(c1, substrate_corner(P1, Ci, Ksi, ksg, L, l0i));
          c2 = add(c2, substrate_corner(P2, Ci, Ksi, ksg, L, l0i));
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
    __syncthreads();
    if (own) {
#pragma unroll
      for (int k = 0; k < DPM3D_MAXDEG; k++) {
        const int e = vfr[k];  // face*3 + corner
        if (e < 0) break;
        F = add(F, mk(l.fcx[e], l.fcy[e], l.fcz[e]));
      }
    }
  }

  // --------------------------------------------------------------- bending
  if ((mask & M_BENDING) && Kbi != 0.0f) {  // block-uniform
    if (own) {
      f3 s = {0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int k = 0; k < DPM3D_MAXDEG; k++) {
        const int u = adjr[k];
        if (u < 0) break;
        s = add(s, sub(lds_v(l.sx, l.sy, l.sz, u), P));
      }
      l.sb[3 * tid] = s.x;
      l.sb[3 * tid + 1] = s.y;
      l.sb[3 * tid + 2] = s.z;
    }
    __syncthreads();
    if (own) {
      f3 acc = {0.0f, 0.0f, 0.0f};
      int deg = 0;
#pragma unroll
      for (int k = 0; k < DPM3D_MAXDEG; k++) {
        const int u = adjr[k];
        if (u < 0) break;
        deg++;
        acc = add(acc, mk(l.sb[3 * u], l.sb[3 * u + 1], l.sb[3 * u + 2]));
      }
      const f3 sv = mk(l.sb[3 * tid], l.sb[3 * tid + 1], l.sb[3 * tid + 2]);
      const f3 lap = sub(acc, scale(sv, (float)deg));
      F = add(F, scale(lap, -Kbi));
    }
  }

  // ------------------------------------------- contact and vertex adhesion
  const bool do_contact = (mask & M_CONTACT) && Kre != 0.0f;
  const bool do_adh = (mask & M_ADHESION) && Kat != 0.0f;
  unsigned long long n_wind = 0, n_contact = 0, n_cand = 0;
  if (do_contact || do_adh) {  // block-uniform
    const float L0i = fmaxf(l0i, 1e-12f);
    float4* ut = l.utab + (size_t)wave * nv;
    for (int c0 = cb; c0 < ce; c0 += NT) {
      const int cj = c0 + tid;
      bool cand = false;
      if (cj < ce && cj != ci) {
        const float4 stj = a.st_in[cj];
        f3 sh = {0.0f, 0.0f, 0.0f};
        if (pbc) {
          sh.x = L * roundf((Ci.x - stj.x) / L);
          sh.y = L * roundf((Ci.y - stj.y) / L);
          sh.z = L * roundf((Ci.z - stj.z) / L);
        }
        const f3 d = mk(stj.x + sh.x - Ci.x, stj.y + sh.y - Ci.y, stj.z + sh.z - Ci.z);
        const float dist = sqrtf(dot(d, d));
        float reach = 0.0f;
        if (do_contact) reach = Ri + stj.w + contact_margin(stj.w, a.em_in[cj]);
        if (do_adh) {
          const float l0m = fmaxf(L0i, fmaxf(a.l0[cj], 1e-12f));
          const float r2 = Ri + stj.w + 2.0f * l0m;
          reach = fmaxf(reach, r2);
          // Adhesion takes the minimum image per vertex pair
          // (Cell3D_Kernel.cl:341-346), which can differ from the
          // centre-of-mass image when the box is small: then no culling.
          if (pbc && 2.0f * r2 >= 0.999f * L) cand = true;
        }
        cand = cand || dist <= reach * 1.0001f + 1e-6f;
      }
      const int ncand = block_compact(cand, cj, l.clist, l.cnt);
      n_cand += ncand;
      for (int q = 0; q < ncand; q++) {
        const int cj2 = l.clist[q];
        const float4 stj = a.st_in[cj2];
        f3 sh = {0.0f, 0.0f, 0.0f};
        if (pbc) {
          sh.x = L * roundf((Ci.x - stj.x) / L);
          sh.y = L * roundf((Ci.y - stj.y) / L);
          sh.z = L * roundf((Ci.z - stj.z) / L);
        }
        const size_t bj = (size_t)cj2 * nv;
        for (int v = tid; v < nv; v += NT) {
          l.nx[v] = a.xin[bj + v];
          l.ny[v] = a.yin[bj + v];
          l.nz[v] = a.zin[bj + v];
        }
        if (tid < DPM3D_MAXCLUS) l.crad[tid] = 0u;
        __syncthreads();
        if (do_adh && own) {
          // radius of each vertex cluster about its seed vertex
          const int k = a.clus_of[tid];
          const f3 d = sub(lds_v(l.nx, l.ny, l.nz, tid), lds_v(l.nx, l.ny, l.nz, k));
          atomicMax(&l.crad[k], __float_as_uint(sqrtf(dot(d, d))));
        }
        __syncthreads();
        if (do_contact) {
          const f3 Cj = mk(stj.x + sh.x, stj.y + sh.y, stj.z + sh.z);
          const float rj = (stj.w + contact_margin(stj.w, a.em_in[cj2])) * 1.0001f + 1e-6f;
          bool need = false;
          if (own) {
            const f3 d = sub(P, Cj);
            need = dot(d, d) <= rj * rj;
          }
          const int np = block_compact(need, tid, l.plist, l.cnt);
          n_wind += np;
          if (wave < a.ww) {
            for (int pi = wave; pi < np; pi += a.ww) {
              const int v = l.plist[pi];
              const float wv = winding_wave(l, ut, nv, nf, sh, lds_v(l.sx, l.sy, l.sz, v));
              if (lane == 0) l.wb[v] = wv;
            }
          }
          __syncthreads();
          if (need) {
            const float wv = l.wb[tid];
            const size_t pidx = (base + tid) * (size_t)a.ncmax + (size_t)(cj2 - cb);
            if (a.p_winding) a.p_winding[pidx] = wv;
            if (!(fabsf(wv) < 1e-6f)) {
              const f3 dir = normalize(sub(Ci, P));
              const float mag = fabsf(wv) * 0.5f * Kre;
              F = add(F, scale(dir, mag));
              n_contact++;
              if (a.p_contact) a.p_contact[pidx] = 1;
            }
          }
        }
        int nact = 0;
        const float L0j = fmaxf(a.l0[cj2], 1e-12f);
        const float l0m = fmaxf(L0i, L0j);
        const bool all_images = pbc && 2.0f * (Ri + stj.w + 2.0f * l0m) >= 0.999f * L;
        if (do_adh) {  // block-uniform
          // clusters of cell j that can reach this cell at all
          bool act = false;
          if (tid < a.nclus) {
            const f3 s = mk(l.nx[tid] + sh.x, l.ny[tid] + sh.y, l.nz[tid] + sh.z);
            const f3 ds = sub(s, Ci);
            const float rk = (Ri + __uint_as_float(l.crad[tid]) + 2.0f * l0m) * 1.0001f + 1e-6f;
            act = all_images || dot(ds, ds) <= rk * rk;
          }
          nact = block_compact(act, tid, l.cact, l.cnt);
        }
        if (do_adh && own && nact > 0) {
          const f3 Cj = mk(stj.x + sh.x, stj.y + sh.y, stj.z + sh.z);
          const f3 d = sub(P, Cj);
          const float r2 = (stj.w + 2.0f * l0m) * 1.0001f + 1e-6f;
          if (all_images || dot(d, d) <= r2 * r2) {
            int cnt_own = 0;
            const float half_l = 0.49f * L;
            const float lim_own = (L0i * 2.0f) * (L0i * 2.0f) * 1.00001f;
            const float lim_rev = (L0j * 2.0f) * (L0j * 2.0f) * 1.00001f;
            for (int q2 = 0; q2 < nact; q2++) {
              const int k = l.cact[q2];
              if (!all_images) {
                // the cluster's vertices lie within crad[k] of its seed
                const f3 s = mk(l.nx[k] + sh.x, l.ny[k] + sh.y, l.nz[k] + sh.z);
                const f3 ds = sub(P, s);
                const float rk = (__uint_as_float(l.crad[k]) + 2.0f * l0m) * 1.0001f + 1e-6f;
                if (dot(ds, ds) > rk * rk) continue;
              }
              for (int m = l.cbeg[k]; m < l.cbeg[k + 1]; m++) {
                const int vj = l.smem_members[m];
                const f3 pj = lds_v(l.nx, l.ny, l.nz, vj);
                // this vertex's own loop: delta = pj - p, minimum image per
                // component (Cell3D_Kernel.cl:341-346); a component below
                // 0.49 L rounds to image 0, which leaves it unchanged.
                f3 dl = sub(pj, P);
                f3 dr = sub(P, pj);  // the partner's loop: delta = p - pj
                f3 nl = {0.0f, 0.0f, 0.0f}, nr = nl;  // image indices
                if (pbc && !(fabsf(dl.x) < half_l && fabsf(dl.y) < half_l && fabsf(dl.z) < half_l)) {
                  nl = mk(roundf(dl.x / L), roundf(dl.y / L), roundf(dl.z / L));
                  nr = mk(roundf(dr.x / L), roundf(dr.y / L), roundf(dr.z / L));
                  dl = sub(dl, mk(L * nl.x, L * nl.y, L * nl.z));
                  dr = sub(dr, mk(L * nr.x, L * nr.y, L * nr.z));
                }
                // sqrt(d2) < c is only possible for d2 < c^2 (1 + 1e-5):
                // the exact comparisons run where they can succeed.  The
                // decisions are the reference's float ones; the force of an
                // accepted pair is evaluated in double on the same image,
                // since across the periodic boundary the float difference
                // of two near-L numbers keeps few digits of a short vector.
                const float d2 = dot(dl, dl);
                if (d2 < lim_own) {
                  const float ds = sqrtf(d2);
                  if (ds < L0i * 2.0f && ds > 1e-12f) {
                    F = sub(F, adhesion_term(pj, P, nl, L, Kat, L0i));
                    cnt_own++;
                  }
                }
                const float d2r = dot(dr, dr);
                if (d2r < lim_rev) {
                  const float dsr = sqrtf(d2r);
                  if (dsr < L0j * 2.0f && dsr > 1e-12f) F = add(F, adhesion_term(P, pj, nr, L, Kat, L0j));
                }
              }
            }
            if (a.p_adh) a.p_adh[base + tid] += cnt_own;
          }
        }
        __syncthreads();
      }
    }
  }

  // ------------------------------------------------------- Euler and stats
  if (own && (a.flags & F_WRITE_FORCES)) {
    a.fx[base + tid] = F.x;
    a.fy[base + tid] = F.y;
    a.fz[base + tid] = F.z;
  }
  if (a.flags & F_MOVE) {  // block-uniform
    f3 Pn = P;
    if (own) {
      Pn = mk(fmaf(F.x, a.dt, P.x), fmaf(F.y, a.dt, P.y), fmaf(F.z, a.dt, P.z));
      a.xout[base + tid] = Pn.x;
      a.yout[base + tid] = Pn.y;
      a.zout[base + tid] = Pn.z;
    }
    __syncthreads();
    if (own) {
      l.sx[tid] = Pn.x;
      l.sy[tid] = Pn.y;
      l.sz[tid] = Pn.z;
    }
    __syncthreads();
    cell_stats(l, nv, nf, a.st_out, a.vol_out, a.em_out, ci);
  }
  if (a.counters) {
    const double c = block_sum((double)n_contact, l.red);
    if (tid == 0) {
      // spread over DPM_COUNTER_SLOTS slots so that workgroups do not
      // serialise on one address; the host sums the slots
      unsigned long long* slot = a.counters + 4 * (blockIdx.x % DPM_COUNTER_SLOTS);
      if (blockIdx.x == 0) atomicAdd(&slot[0], 1ull);
      if (n_wind) atomicAdd(&slot[1], n_wind);
      if (c != 0.0) atomicAdd(&slot[2], (unsigned long long)c);
      if (n_cand) atomicAdd(&slot[3], n_cand);
    }
  }
}
