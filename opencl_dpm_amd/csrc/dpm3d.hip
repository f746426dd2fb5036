// 3D DPM timestep kernels for gfx950.
//
// One workgroup owns one cell for a whole step (grid = cells of all tissues,
// workgroups spread over the 8 XCDs so that neighbouring cells share an L2).
// The cell's vertices live in LDS for the step.  Every force term of the
// reference's per-face kernels is evaluated by *gathering* over the faces
// incident to a vertex in a fixed order, so no float atomics are used and
// a step is bit-reproducible run to run (the reference scatters with a
// compare-and-swap float add, Cell3D_Kernel.cl:10-32, whose order varies).
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
  float4* st_out;
  double* vol_out;
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
  const int* faces;  // [nf*3]
  const int* vf;     // [nv*MAXDEG] incident (face*3 + corner), ascending, -1 pad
  const int* adj;    // [nv*MAXDEG] neighbour vertices, ascending, -1 pad
  int nv, nf, ncells;
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
// test.  Outside radius R_j + m_j of the (shifted) centre of mass every face
// of cell j is at least m_j away, so with m_j >= 2 l0 no face can subtend the
// more-than-hemisphere solid angle the reference skips (denom < 1e-8), the
// exact winding number is 0 and its float evaluation stays orders of
// magnitude below the 1e-6 pair threshold.
__device__ __forceinline__ float contact_margin(float Rj, float l0j) {
  return 2.0f * l0j + 0.1f * Rj;
}

// LDS layout shared by both kernels (sizes in elements):
//   float own[3*nv] | float nbr[3*nv] | float bend[3*nv] | float wnum[nv]
//   int clist[NT] | int plist[NT] | int cnt[32] | (8-aligned) double red[32]
__host__ __device__ inline size_t lds3d_bytes(int nv, int nt) {
  size_t b = sizeof(float) * 10u * (size_t)nv + sizeof(int) * (2u * (size_t)nt + 32u);
  b = (b + 7u) & ~(size_t)7u;
  return b + sizeof(double) * 32u;
}

struct Lds3D {
  float *sx, *sy, *sz, *nx, *ny, *nz, *sb, *wb;
  int *clist, *plist, *cnt;
  double* red;
};
__device__ __forceinline__ Lds3D carve3d(unsigned char* smem, int nv, int nt) {
  Lds3D l;
  float* f = reinterpret_cast<float*>(smem);
  l.sx = f;
  l.sy = f + nv;
  l.sz = f + 2 * nv;
  l.nx = f + 3 * nv;
  l.ny = f + 4 * nv;
  l.nz = f + 5 * nv;
  l.sb = f + 6 * nv;
  l.wb = f + 9 * nv;
  int* ip = reinterpret_cast<int*>(f + 10 * nv);
  l.clist = ip;
  l.plist = ip + nt;
  l.cnt = ip + 2 * nt;
  size_t off = sizeof(float) * 10u * (size_t)nv + sizeof(int) * (2u * (size_t)nt + 32u);
  off = (off + 7u) & ~(size_t)7u;
  l.red = reinterpret_cast<double*>(smem + off);
  return l;
}

// Centre of mass (double sums), bounding radius and volume of the cell whose
// vertices are in l.sx/sy/sz.  The volume is summed over faces relative to
// the centre of mass (translation-invariant for a closed surface), per-face
// terms in float, the sum in double — the reference sums absolute-coordinate
// triple products in float (getVolume, Cell3D_Kernel.cl:46-64), which
// cancels catastrophically far from the origin.
__device__ void cell_stats(const Lds3D& l, const int* __restrict__ faces, int nv, int nf,
                           float4* st_out, double* vol_out, int ci) {
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
    f3 d = sub(mk(l.sx[v], l.sy[v], l.sz[v]), C);
    r = fmaxf(r, sqrtf(dot(d, d)));
  }
  r = block_max(r, reinterpret_cast<float*>(l.red));
  double vol = 0.0;
  for (int f = tid; f < nf; f += NT) {
    const int i0 = faces[3 * f], i1 = faces[3 * f + 1], i2 = faces[3 * f + 2];
    f3 P0 = sub(mk(l.sx[i0], l.sy[i0], l.sz[i0]), C);
    f3 P1 = sub(mk(l.sx[i1], l.sy[i1], l.sz[i1]), C);
    f3 P2 = sub(mk(l.sx[i2], l.sy[i2], l.sz[i2]), C);
    vol += (double)dot(cross(P0, P1), P2);
  }
  vol = block_sum(vol, l.red);
  if (tid == 0) {
    // Slack of a few ulp so that the culling tests are never tighter than the
    // largest float distance a vertex can have from C.
    st_out[ci] = make_float4(C.x, C.y, C.z, r * 1.0000005f + 1e-30f);
    vol_out[ci] = fabs(vol / 6.0);
  }
}

extern "C" __global__ void __launch_bounds__(1024) k3d_stats(const Step3D a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int nv = a.nv, NT = blockDim.x, tid = threadIdx.x;
  const int ci = xcd_remap(blockIdx.x, gridDim.x);
  Lds3D l = carve3d(smem, nv, NT);
  const size_t base = (size_t)ci * nv;
  for (int v = tid; v < nv; v += NT) {
    l.sx[v] = a.xin[base + v];
    l.sy[v] = a.yin[base + v];
    l.sz[v] = a.zin[base + v];
  }
  __syncthreads();
  cell_stats(l, a.faces, nv, a.nf, a.st_out, a.vol_out, ci);
}

// One overdamped Euler step of every cell (or, without F_MOVE, one force
// evaluation for diagnostics).  blockDim.x >= nv, a multiple of 64.
extern "C" __global__ void __launch_bounds__(1024) k3d_step(const Step3D a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int nv = a.nv, nf = a.nf, NT = blockDim.x, tid = threadIdx.x;
  const int lane = tid & (DPM_WAVE - 1), wave = tid / DPM_WAVE, nw = NT / DPM_WAVE;
  const int ci = xcd_remap(blockIdx.x, gridDim.x);
  Lds3D l = carve3d(smem, nv, NT);

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

  for (int v = tid; v < nv; v += NT) {
    l.sx[v] = a.xin[base + v];
    l.sy[v] = a.yin[base + v];
    l.sz[v] = a.zin[base + v];
  }
  __syncthreads();

  const bool own = tid < nv;
  const f3 P = own ? mk(l.sx[tid], l.sy[tid], l.sz[tid]) : mk(0.0f, 0.0f, 0.0f);
  f3 F = {0.0f, 0.0f, 0.0f};

  // ---------------------------------------------------------- shape forces
  if (own) {
    // Volume strain of the current volume, in double (the residual V/v0 - 1
    // cancels when the cell is near its target).
    const float strain = (float)(a.vol_in[ci] / (double)v0i - 1.0);
    const float prevol = -Kvi * strain;
    const float sc = Kai * sqrtf(a0i) / l0i * 0.3f;
    for (int k = 0; k < DPM3D_MAXDEG; k++) {
      const int e = a.vf[tid * DPM3D_MAXDEG + k];
      if (e < 0) break;
      const int f = e / 3, corner = e - 3 * f;
      const int i0 = a.faces[3 * f], i1 = a.faces[3 * f + 1], i2 = a.faces[3 * f + 2];
      const f3 P0 = mk(l.sx[i0], l.sy[i0], l.sz[i0]);
      const f3 P1 = mk(l.sx[i1], l.sy[i1], l.sz[i1]);
      const f3 P2 = mk(l.sx[i2], l.sy[i2], l.sz[i2]);
      if ((mask & M_VOLUME) && Kvi != 0.0f) {
        const f3 A = sub(P1, Ci), B = sub(P2, Ci), Cc = sub(P0, Ci);
        const f3 g = corner == 0 ? cross(A, B) : (corner == 1 ? cross(B, Cc) : cross(Cc, A));
        F.x += prevol * g.x / 6.0f;
        F.y += prevol * g.y / 6.0f;
        F.z += prevol * g.z / 6.0f;
      }
      if ((mask & M_SURFACE) && !(Kai < 1e-8f)) {
        const f3 lv0 = sub(P1, P0), lv1 = sub(P2, P1), lv2 = sub(P0, P2);
        const float len0 = sqrtf(dot(lv0, lv0)), len1 = sqrtf(dot(lv1, lv1)),
                    len2 = sqrtf(dot(lv2, lv2));
        if (!(len0 < 1e-12f || len1 < 1e-12f || len2 < 1e-12f)) {
          const f3 u0 = mk(lv0.x / len0, lv0.y / len0, lv0.z / len0);
          const f3 u1 = mk(lv1.x / len1, lv1.y / len1, lv1.z / len1);
          const f3 u2 = mk(lv2.x / len2, lv2.y / len2, lv2.z / len2);
          const float d0 = (len0 / l0i) - 1.0f, d1 = (len1 / l0i) - 1.0f, d2 = (len2 / l0i) - 1.0f;
          f3 g;
          if (corner == 0)
            g = sub(scale(u0, d0), scale(u2, d2));
          else if (corner == 1)
            g = sub(scale(u1, d1), scale(u0, d0));
          else
            g = sub(scale(u2, d2), scale(u1, d1));
          F = add(F, scale(g, sc));
        }
      }
      if ((mask & M_SUBSTRATE) && !(Ksi < 1e-12f)) {
        const f3 un = normalize(cross(sub(P1, P0), sub(P2, P0)));
        if (un.z < -0.1f) {
          if (a.p_down && corner == 0) a.p_down[(size_t)ci * nf + f] = 1;
          const f3 pos = P;  // this vertex is corner `corner` of face f
          float ks = Ksi;
          if (ksg != 0.0f) {
            const float xw = pos.x - L * floorf(pos.x / L);
            ks = Ksi * (1.0f + ksg * xw / L);
          }
          if (pos.z < 0.0f) F.z += ks * fabsf(pos.z);
          const float height = fabsf(pos.z);
          if (height < l0i * 2.0f) {
            const f3 c2v = mk(pos.x - Ci.x, pos.y - Ci.y, 0.0f - Ci.z);
            const float dz = pos.z - 0.0f;
            const float dist = sqrtf((0.0f * 0.0f + 0.0f * 0.0f) + dz * dz);
            const float ftmp = (float)((double)ks * (1.0 - (double)(dist / l0i)));
            F = add(F, scale(normalize(c2v), ftmp));
          }
        }
      }
    }
  }

  // --------------------------------------------------------------- bending
  if ((mask & M_BENDING) && Kbi != 0.0f) {  // block-uniform
    if (own) {
      f3 s = {0.0f, 0.0f, 0.0f};
      for (int k = 0; k < DPM3D_MAXDEG; k++) {
        const int u = a.adj[tid * DPM3D_MAXDEG + k];
        if (u < 0) break;
        s = add(s, sub(mk(l.sx[u], l.sy[u], l.sz[u]), P));
      }
      l.sb[3 * tid] = s.x;
      l.sb[3 * tid + 1] = s.y;
      l.sb[3 * tid + 2] = s.z;
    }
    __syncthreads();
    if (own) {
      f3 acc = {0.0f, 0.0f, 0.0f};
      int deg = 0;
      for (int k = 0; k < DPM3D_MAXDEG; k++) {
        const int u = a.adj[tid * DPM3D_MAXDEG + k];
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
        if (do_contact) reach = Ri + stj.w + contact_margin(stj.w, a.l0[cj]);
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
        __syncthreads();
        if (do_contact) {
          const f3 Cj = mk(stj.x + sh.x, stj.y + sh.y, stj.z + sh.z);
          const float rj = (stj.w + contact_margin(stj.w, a.l0[cj2])) * 1.0001f + 1e-6f;
          bool need = false;
          if (own) {
            const f3 d = sub(P, Cj);
            need = dot(d, d) <= rj * rj;
          }
          const int np = block_compact(need, tid, l.plist, l.cnt);
          n_wind += np;
          for (int pi = wave; pi < np; pi += nw) {
            const int v = l.plist[pi];
            const f3 p = mk(l.sx[v], l.sy[v], l.sz[v]);
            float part = 0.0f;
            for (int f = lane; f < nf; f += DPM_WAVE) {
              const int ia = a.faces[3 * f], ib = a.faces[3 * f + 1], ic = a.faces[3 * f + 2];
              // Verts + shift - point, in the reference's order.
              const f3 A = mk((l.nx[ia] + sh.x) - p.x, (l.ny[ia] + sh.y) - p.y, (l.nz[ia] + sh.z) - p.z);
              const f3 B = mk((l.nx[ib] + sh.x) - p.x, (l.ny[ib] + sh.y) - p.y, (l.nz[ib] + sh.z) - p.z);
              const f3 C = mk((l.nx[ic] + sh.x) - p.x, (l.ny[ic] + sh.y) - p.y, (l.nz[ic] + sh.z) - p.z);
              const f3 u = normalize(A), vv = normalize(B), w = normalize(C);
              const float denom = ((1.0f + dot(u, vv)) + dot(vv, w)) + dot(w, u);
              if (denom < 1e-8f) continue;
              const float num = dot(u, cross(vv, w));
              part += 2.0f * atan2f(num, denom);
            }
            const double tot = wave_sum((double)part);
            if (lane == 0) l.wb[v] = (float)(tot / (double)(4.0f * 3.14159265358979323846f));
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
        if (do_adh && own) {
          const float L0j = fmaxf(a.l0[cj2], 1e-12f);
          const f3 Cj = mk(stj.x + sh.x, stj.y + sh.y, stj.z + sh.z);
          const f3 d = sub(P, Cj);
          const float r2 = (stj.w + 2.0f * fmaxf(L0i, L0j)) * 1.0001f + 1e-6f;
          const bool all_images = pbc && 2.0f * (Ri + stj.w + 2.0f * fmaxf(L0i, L0j)) >= 0.999f * L;
          if (all_images || dot(d, d) <= r2 * r2) {
            int cnt_own = 0;
            for (int vj = 0; vj < nv; vj++) {
              const f3 pj = mk(l.nx[vj], l.ny[vj], l.nz[vj]);
              // this vertex's own loop: delta = pj - p
              f3 dl = sub(pj, P);
              if (pbc) {
                const float rx = roundf(dl.x / L), ry = roundf(dl.y / L), rz = roundf(dl.z / L);
                dl = sub(dl, mk(L * rx, L * ry, L * rz));
              }
              const float ds = sqrtf(dot(dl, dl));
              if (ds < L0i * 2.0f && ds > 1e-12f) {
                const float k = Kat * 0.5f * (ds / L0i - 1.0f);
                F = sub(F, mk(k * (dl.x / ds), k * (dl.y / ds), k * (dl.z / ds)));
                cnt_own++;
              }
              // the partner's loop applies +t to this vertex: delta = p - pj
              f3 dr = sub(P, pj);
              if (pbc) {
                const float rx = roundf(dr.x / L), ry = roundf(dr.y / L), rz = roundf(dr.z / L);
                dr = sub(dr, mk(L * rx, L * ry, L * rz));
              }
              const float dsr = sqrtf(dot(dr, dr));
              if (dsr < L0j * 2.0f && dsr > 1e-12f) {
                const float k = Kat * 0.5f * (dsr / L0j - 1.0f);
                F = add(F, mk(k * (dr.x / dsr), k * (dr.y / dsr), k * (dr.z / dsr)));
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
    cell_stats(l, a.faces, nv, nf, a.st_out, a.vol_out, ci);
  }
  if (a.counters) {
    // one atomic per workgroup per counter (each thread's copy is identical
    // for the block-level counts; per-vertex counts are summed first).
    double c = block_sum((double)n_contact, l.red);
    if (tid == 0) {
      if (blockIdx.x == 0) atomicAdd(&a.counters[0], 1ull);
      atomicAdd(&a.counters[1], n_wind);
      atomicAdd(&a.counters[2], (unsigned long long)c);
      atomicAdd(&a.counters[3], n_cand);
    }
  }
}
