// Wave-per-cell form of the 3D step (one 64-lane wave owns one cell).
//
// The same step as k3d_step (dpm3d.hip) for meshes of at most 256
// vertices: lane l handles vertices l, l + 64, ..., the per-vertex state
// (positions, force accumulators) lives in the wave's LDS slice, the
// incidence and neighbour rows are read as 16-byte words of an L1-resident
// topology image, every per-cell reduction is a wave shuffle and every list
// a wave ballot, and the waves of a workgroup never wait for each other
// after the faces are staged.  The LDS slice and a 128-register budget are
// sized for four waves per SIMD.  Shape forces are gathered per vertex: volume and substrate terms
// per incident face, edge springs once per ring edge (each lies on two of
// the vertex's faces), so the sums are grouped differently from the face
// pass of k3d_step (float rounding); bending and contact use the same
// arithmetic; the adhesion sums are grouped in slices of one wave.
#pragma once

constexpr int kWaveGlist = 128;  // gathered cell-list candidates per wave
constexpr int kWaveMaxNv = 256;  // meshes the wave-per-cell step takes
constexpr int kWaveCand = 128;   // candidate cells per batch
static_assert(kWaveGlist <= kWaveCand, "the gathered list is sorted into the candidate array");

// LDS: packed faces shared by the workgroup, then one slice per wave.  The
// winding directions (or adhesion slices) share their bytes with the
// cell-list gather, which is finished before the first candidate, and with
// the per-vertex facing-down counts of the shape phase, finished before the
// candidate search.
struct Lds3DWave {
  size_t vf, adj, img, fend, slice, total;
  size_t sx, nx, fx, il, down, cl, ut, gl, gath;
};
__host__ __device__ inline Lds3DWave ldsw_layout(int nv, int nf, int cpw, int rsize) {
  Lds3DWave L;
  const size_t rs = (size_t)rsize, n = (size_t)nv;
  auto al = [](size_t o) { return (o + 31u) & ~(size_t)31u; };
  // the topology image (global, L1/L2-resident): packed faces, then the
  // incidence and neighbour rows, one 16-byte row of 16-bit entries per vertex
  L.fend = al(4u * (size_t)nf);          // faces: i0 | i1 << 10 | i2 << 20 (also in LDS)
  L.vf = L.fend;                          // incident (face*3 + corner) rows (0xffff: none)
  L.adj = L.vf + 2u * n * DPM3D_MAXDEG;   // neighbour rows
  L.img = al(L.adj + 2u * n * DPM3D_MAXDEG);
  size_t s = 0;
  L.sx = s;   s = al(s + 3u * rs * n);  // own positions (x | y | z)
  L.nx = s;   s = al(s + 3u * rs * n);  // candidate positions (those in reach compacted in place); bending scratch
  L.fx = s;   s = al(s + 3u * rs * n);  // force accumulators
  L.il = s;   s = al(s + 2u * n);       // own vertices in reach
  L.down = s; s = al(s + 4u * (size_t)((nf + 31) / 32));  // facing-down faces
  L.cl = s;   s = al(s + 4u * (size_t)kWaveCand);  // candidate cells of a batch
  L.ut = s;
  const size_t ut_end = al(s + 4u * rs * (n > 64 ? n : 64));  // winding directions | adhesion slices
  L.gl = s;   s += 4u * (size_t)kWaveGlist;  // cell-list candidates
  L.gath = s; s += 4u * 96u;                  // 27 buckets: ids, starts, prefix
  s = al(s);
  L.slice = s > ut_end ? s : ut_end;
  L.total = L.fend + (size_t)cpw * L.slice;
  return L;
}

// Entry q of a 16-byte row of eight 16-bit entries.
__device__ __forceinline__ int row16(uint4 r, int q) {
  const unsigned w = q < 4 ? (q < 2 ? r.x : r.y) : (q < 6 ? r.z : r.w);
  return (int)((w >> ((q & 1) * 16)) & 0xffffu);
}

__device__ __forceinline__ void wface(const unsigned* sf, int f, int& i0, int& i1, int& i2) {
  const unsigned w = sf[f];
  i0 = (int)(w & 1023u);
  i1 = (int)((w >> 10) & 1023u);
  i2 = (int)(w >> 20);
}

// Wave sums: every lane receives the result (DPP tree, fixed order).
__device__ __forceinline__ double wsum_all(double v) { return wave_sum_d(v); }

// Set bits of a ballot in the lanes below this one (v_mbcnt).
__device__ __forceinline__ int lanes_below(unsigned long long m) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// A wave-uniform value computed by vector instructions, moved to a scalar
// register (every lane holds the same bits).
__device__ __forceinline__ float wave_uniform(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ __forceinline__ double wave_uniform(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// One cell's positions (nv <= kWaveMaxNv) into a wave's LDS arrays, every
// load issued before the first store (one memory round trip, not one per
// 64 vertices).
template <typename R>
__device__ __forceinline__ void wave_stage_cell(R* dx, R* dy, R* dz, const R* x, const R* y, const R* z,
                                                size_t base, int nv, int lane) {
  R bx[kWaveMaxNv / DPM_WAVE], by[kWaveMaxNv / DPM_WAVE], bz[kWaveMaxNv / DPM_WAVE];
#pragma unroll
  for (int k = 0; k < kWaveMaxNv / DPM_WAVE; k++) {
    const int v = lane + k * DPM_WAVE;
    if (v < nv) {
      bx[k] = x[base + v];
      by[k] = y[base + v];
      bz[k] = z[base + v];
    }
  }
#pragma unroll
  for (int k = 0; k < kWaveMaxNv / DPM_WAVE; k++) {
    const int v = lane + k * DPM_WAVE;
    if (v < nv) {
      dx[v] = bx[k];
      dy[v] = by[k];
      dz[v] = bz[k];
    }
  }
}

// winding_wave over per-wave arrays and packed faces (same arithmetic).
template <typename R>
__device__ __forceinline__ R winding_w(const R* nx, const R* ny, const R* nz, const unsigned* sf, Q4<R>* ut,
                                       int nv, int nf, V3<R> sh, V3<R> p) {
  const int lane = threadIdx.x & (DPM_WAVE - 1);
  #pragma unroll 1
  for (int v = lane; v < nv; v += DPM_WAVE) {
    const V3<R> A = {(nx[v] + sh.x) - p.x, (ny[v] + sh.y) - p.y, (nz[v] + sh.z) - p.z};
    const V3<R> u = vnormalize(A);
    ut[v] = Q4<R>{u.x, u.y, u.z, R(0)};
  }
  wave_lds_sync();
  R part = R(0);
  #pragma unroll 1
  for (int f = lane; f < nf; f += DPM_WAVE) {
    int ia, ib, ic;
    wface(sf, f, ia, ib, ic);
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
  wave_lds_sync();
  return (R)(tot / (double)(R(4) * R(3.14159265358979323846)));
}

// NVC, NFC: vertex and face counts fixed at compile time (0: taken from the
// arguments).  Fixed sizes turn every per-wave LDS array offset into an
// immediate (the 162-vertex icosphere).  The vertex and face loops stay
// rolled: unrolled, their iterations' registers spill to scratch memory
// (208 B per lane, 6x the step's HBM traffic, at the same speed).
template <typename R, int NVC, int NFC, bool PROBE = true>
__device__ void k3d_wstep_body(const Step3D<R>& a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int nv = NVC > 0 ? NVC : a.nv, nf = NFC > 0 ? NFC : a.nf;
  // 64-vertex chunks of a cell (a bound when the size is a run-time value)
  constexpr int kChunks = NVC > 0 ? (NVC + DPM_WAVE - 1) / DPM_WAVE : kWaveMaxNv / DPM_WAVE;
  // wv, and every per-cell value derived from it, is wave-uniform: through
  // readfirstlane it lives in scalar registers (the cell's parameters and
  // statistics come by scalar loads), which keeps the step within 128 vector
  // registers without spilling to scratch memory
  const int lane = threadIdx.x & (DPM_WAVE - 1), wv = __builtin_amdgcn_readfirstlane(threadIdx.x / DPM_WAVE);
  const int cpw = blockDim.x / DPM_WAVE;
  const Lds3DWave LW = ldsw_layout(nv, nf, cpw, (int)sizeof(R));
  unsigned* sf = reinterpret_cast<unsigned*>(smem);
  const uint4* rows = a.wtopo;  // 16-byte words of the topology image
  for (int i = threadIdx.x; i < (int)(LW.fend / 16); i += blockDim.x) reinterpret_cast<uint4*>(smem)[i] = rows[i];
  __syncthreads();  // the only workgroup barrier
  const int ci = xcd_remap(blockIdx.x, gridDim.x) * cpw + wv;
  if (ci >= a.ncells) return;  // whole wave

  unsigned char* slb = smem + LW.fend + (size_t)wv * LW.slice;
  R* sx = reinterpret_cast<R*>(slb + LW.sx);
  R* sy = sx + nv;
  R* sz = sy + nv;
  R* nx = reinterpret_cast<R*>(slb + LW.nx);
  R* ny = nx + nv;
  R* nz = ny + nv;
  R* fx = reinterpret_cast<R*>(slb + LW.fx);
  R* fy = fx + nv;
  R* fz = fy + nv;
  unsigned short* il = reinterpret_cast<unsigned short*>(slb + LW.il);
  Q4<R>* ut = reinterpret_cast<Q4<R>*>(slb + LW.ut);
  int* gl = reinterpret_cast<int*>(slb + LW.gl);
  int* gath = reinterpret_cast<int*>(slb + LW.gath);
  int* cl = reinterpret_cast<int*>(slb + LW.cl);
  unsigned* down = reinterpret_cast<unsigned*>(slb + LW.down);

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
  const V3<R> zero = {R(0), R(0), R(0)};

  wave_stage_cell(sx, sy, sz, a.xin, a.yin, a.zin, base, nv, lane);
  #pragma unroll 1
  for (int v = lane; v < nv; v += DPM_WAVE) {
    if (a.w_out) a.w_out[base + v] = R(0);
    if (a.rec_cell)
      for (int q = 0; q < DPM3D_MAXPART; q++) a.rec_cell[(base + v) * DPM3D_MAXPART + q] = -1;
  }
  wave_lds_sync();

  // ---------------------------------------------------------- shape forces
  const bool shape_on = (mask & (M_VOLUME | M_SURFACE | M_SUBSTRATE)) != 0;
  const R strain = (R)(a.vol_in[ci] / (double)v0i - 1.0);
  const R kvol = -Kvi * strain / R(6);
  const bool vol_on = (mask & M_VOLUME) && Kvi != R(0);
  // the ring walk sums the volume products unconditionally and scales them
  // by 0 when the term is off (a finite sum times 0 adds a signed zero)
  const R kvol_on = vol_on ? kvol : R(0);
  const bool surf_on = (mask & M_SURFACE) && !(Kai < R(1e-8f));
  const bool sub_on = (mask & M_SUBSTRATE) && !(Ksi < R(1e-12f));
  const R sc = Kai * sqrt(a0i) / l0i * R(0.3f);
  const R inv_l0 = R(1) / l0i;
  const double sub_inv_l0 = 1.0 / (double)l0i;
  // Face pre-pass: the facing-down faces (Cell3D_Kernel.cl:206-246), one bit
  // each, and per vertex the number of facing-down faces it is a corner of
  // (integer LDS atomics: the counts do not depend on their order; the array
  // is free until the candidates).
  unsigned* dcnt = reinterpret_cast<unsigned*>(slb + LW.ut);
  if (shape_on && sub_on) {  // wave-uniform
    #pragma unroll 1
    for (int v = lane; v < nv; v += DPM_WAVE) dcnt[v] = 0u;
    wave_lds_sync();
    #pragma unroll 1
    for (int f0 = 0; f0 < nf; f0 += DPM_WAVE) {
      const int f = f0 + lane;
      bool dn = false;
      if (f < nf) {
        int i0, i1, i2;
        wface(sf, f, i0, i1, i2);
        const V3<R> P0 = lds_v(sx, sy, sz, i0), P1 = lds_v(sx, sy, sz, i1), P2 = lds_v(sx, sy, sz, i2);
        dn = facing_down(vcross(vsub(P1, P0), vsub(P2, P0)));
        if (dn) {
          atomicAdd(&dcnt[i0], 1u);
          atomicAdd(&dcnt[i1], 1u);
          atomicAdd(&dcnt[i2], 1u);
        }
        if (PROBE && dn && a.p_down) a.p_down[(size_t)ci * nf + f] = 1;
      }
      const unsigned long long m = __ballot(dn);
      if (lane == 0) {
        down[f0 / 32] = (unsigned)m;
        if (f0 / 32 + 1 < (nf + 31) / 32) down[f0 / 32 + 1] = (unsigned)(m >> 32);
      }
    }
    wave_lds_sync();
  }
  const bool bend_on = (mask & M_BENDING) && Kbi != R(0);
  // The ring walk, unless an edge is shorter than 2e-12 (an edge below 1e-12
  // makes the reference skip its faces' edge springs): the walk checks every
  // ring edge, and a cell with such an edge is taken again, every vertex by
  // the per-corner form (its springs skip those faces as the reference does).
  bool ring_pass = shape_on && a.closed;  // wave-uniform
  for (int pass = 0; pass < 2; pass++) {
  bool short_edge = false;
  #pragma unroll 1
  for (int v = lane; v < nv; v += DPM_WAVE) {
    const V3<R> P = lds_v(sx, sy, sz, v);
    V3<R> F = zero;
    if (ring_pass) {
      // One walk around the vertex's ring u_0 .. u_(k-1), listed so that
      // face q is (v, u_q, u_(q+1)) in its own corner order (closed mesh,
      // every vertex's faces one cycle).  With d_q = u_q - p:
      // volume (Cell3D_Kernel.cl:95-111): corner v of face q takes
      //   kvol (u_q - C) x (u_(q+1) - C); summed over the closed ring this is
      //   kvol sum_q d_q x d_(q+1) (the C terms cancel), taken in that form;
      // edge springs (Cell3D_Kernel.cl:140-176): corner c of a face takes
      //   sc (e_c - e_(c-1)) = sc (s(a - p) + s(b - p)) with
      //   s(d) = d (|d| / l0 - 1) / |d| and a, b its other corners, and every
      //   ring edge lies on two of the vertex's faces: 2 sc sum_q s(d_q);
      // bending's first sum, sum_q d_q, kept in nx for its second pass.
      // (Force arithmetic only, no cut-off: the products fuse into FMAs.)
      const uint4 arow = rows[LW.adj / 16 + v];
      const V3<R> d0 = vsub(lds_v(sx, sy, sz, row16(arow, 0)), P);
      V3<R> dq = d0, Fv = zero, Fs = zero, sb = zero;
      for (int q = 0; q < DPM3D_MAXDEG; q++) {
        if (surf_on) {
          const R d2 = vdot_f(dq, dq);
          short_edge = short_edge || d2 < R(4e-24f);
          Fs = vmadd_f(dq, spring_factor(d2, inv_l0), Fs);
        }
        sb = vadd(sb, dq);
        const int un = q + 1 < DPM3D_MAXDEG ? row16(arow, q + 1) : 0xffff;
        const V3<R> dn = un == 0xffff ? d0 : vsub(lds_v(sx, sy, sz, un), P);
        Fv = vcross_add_f(dq, dn, Fv);
        if (un == 0xffff) break;
        dq = dn;
      }
      F = vadd(vscale(Fv, kvol_on), vscale(Fs, R(2) * sc));
      if (bend_on) {
        nx[v] = sb.x;
        ny[v] = sb.y;
        nz[v] = sb.z;
      }
      if (sub_on) {  // Cell3D_Kernel.cl:206-246, per facing-down incident face
        const unsigned ndown = dcnt[v];
        if (ndown) F = vadd(F, vscale(substrate_corner(P, Ci, Ksi, ksg, L, l0i, sub_inv_l0), (R)ndown));
      }
    } else if (shape_on) {
      const uint4 vrow = rows[LW.vf / 16 + v];
      for (int q = 0; q < DPM3D_MAXDEG; q++) {
        const int e = row16(vrow, q);
        if (e == 0xffff) break;
        const int f = e / 3, c = e - 3 * f;
        int i0, i1, i2;
        wface(sf, f, i0, i1, i2);
        const V3<R> P0 = lds_v(sx, sy, sz, i0), P1 = lds_v(sx, sy, sz, i1), P2 = lds_v(sx, sy, sz, i2);
        // the face pass's corner term for this corner (k3d_step)
        V3<R> cc = zero;
        if (vol_on) {
          const V3<R> A = vsub(P1, Ci), B = vsub(P2, Ci), Cc = vsub(P0, Ci);
          cc = c == 0 ? vscale(vcross(A, B), kvol) : c == 1 ? vscale(vcross(B, Cc), kvol)
                                                            : vscale(vcross(Cc, A), kvol);
        }
        if (surf_on) {
          const V3<R> lv0 = vsub(P1, P0), lv1 = vsub(P2, P1), lv2 = vsub(P0, P2);
          const R len0 = sqrt(vdot(lv0, lv0)), len1 = sqrt(vdot(lv1, lv1)), len2 = sqrt(vdot(lv2, lv2));
          if (!(len0 < R(1e-12f) || len1 < R(1e-12f) || len2 < R(1e-12f))) {
            const V3<R> la = c == 0 ? lv0 : c == 1 ? lv1 : lv2;
            const V3<R> lb = c == 0 ? lv2 : c == 1 ? lv0 : lv1;
            const R na = c == 0 ? len0 : c == 1 ? len1 : len2;
            const R nb = c == 0 ? len2 : c == 1 ? len0 : len1;
            const V3<R> ea = vscale(la, (na * inv_l0 - R(1)) / na);
            const V3<R> eb = vscale(lb, (nb * inv_l0 - R(1)) / nb);
            cc = vadd(cc, vscale(vsub(ea, eb), sc));
          }
        }
        if (sub_on && ((down[f >> 5] >> (f & 31)) & 1u)) cc = vadd(cc, substrate_corner(P, Ci, Ksi, ksg, L, l0i, sub_inv_l0));
        F = vadd(F, cc);
      }
    }
    fx[v] = F.x;
    fy[v] = F.y;
    fz[v] = F.z;
  }
  if (!(ring_pass && surf_on && __ballot(short_edge) != 0ull)) break;  // wave-uniform
  ring_pass = false;
  }

  // --------------------------------------------------------------- bending
  if (bend_on) {  // wave-uniform; scratch in nx
    for (int v = lane; v < nv && !ring_pass; v += DPM_WAVE) {
      const V3<R> P = lds_v(sx, sy, sz, v);
      V3<R> s = zero;
      const uint4 arow = rows[LW.adj / 16 + v];
      for (int q = 0; q < DPM3D_MAXDEG; q++) {
        const int u = row16(arow, q);
        if (u == 0xffff) break;
        s = vadd(s, vsub(lds_v(sx, sy, sz, u), P));
      }
      nx[v] = s.x;
      ny[v] = s.y;
      nz[v] = s.z;
    }
    wave_lds_sync();
    #pragma unroll 1
    for (int v = lane; v < nv; v += DPM_WAVE) {
      V3<R> acc = zero;
      int deg = 0;
      const uint4 arow = rows[LW.adj / 16 + v];
      for (int q = 0; q < DPM3D_MAXDEG; q++) {
        const int u = row16(arow, q);
        if (u == 0xffff) break;
        deg++;
        acc = vadd(acc, lds_v(nx, ny, nz, u));
      }
      const V3<R> lap = vsub(acc, vscale(lds_v(nx, ny, nz, v), (R)deg));
      const V3<R> F = vadd(lds_v(fx, fy, fz, v), vscale(lap, -Kbi));
      fx[v] = F.x;
      fy[v] = F.y;
      fz[v] = F.z;
    }
  }
  wave_lds_sync();  // forces and scratch settled before the candidates

  // ------------------------------------------- contact and vertex adhesion
  const bool do_contact = (mask & M_CONTACT) && Kre != R(0);
  const bool do_adh = (mask & M_ADHESION) && Kat != R(0);
  unsigned n_wind = 0, n_contact = 0, n_cand = 0;  // per step and cell: 32 bits suffice
  if (do_contact || do_adh) {  // wave-uniform
    const R L0i = wave_uniform(fmax(l0i, R(1e-12f)));
    const R invL = wave_uniform(R(1) / L);
    // candidate source: every cell of the tissue, or the cells of the 27
    // buckets around this cell sorted by index (held two per lane)
    bool gathered = false;
    int nsrc = ce - cb;
    if (a.bin_list && a.tbin_nb[t] > 0) {
      const int nb = a.tbin_nb[t], nbuck = a.tbin_nbuck[t];
      const int* bs = a.bin_start + a.tbin_hoff[t];
      const R wb = L / (R)nb;
      int* gst = gath + 32;
      int* gof = gath + 64;
      int g = 0;
      if (lane < 27) {
        const int bx = bin_coord(Ci.x, L, wb, nb), by = bin_coord(Ci.y, L, wb, nb), bz = bin_coord(Ci.z, L, wb, nb);
        const int dx = lane % 3 - 1, dy = (lane / 3) % 3 - 1, dz = lane / 9 - 1;
        g = (int)bin_bucket((bx + dx + nb) % nb, (by + dy + nb) % nb, (bz + dz + nb) % nb, nb, nbuck);
      }
      // distinct buckets, their cells' offsets
      const int nu = distinct_buckets27(g, bs, gst, gof);
      wave_lds_sync();
      const int n = gof[nu];
      if (n <= kWaveGlist) {  // wave-uniform; otherwise scan every cell
        int val0 = -1, val1 = -1;
        if (lane < n) {
          int u = 0;
          while (gof[u + 1] <= lane) u++;
          val0 = a.bin_list[cb + gst[u] + (lane - gof[u])];
          gl[lane] = val0;
        }
        if (lane + DPM_WAVE < n) {
          const int i = lane + DPM_WAVE;
          int u = 0;
          while (gof[u + 1] <= i) u++;
          val1 = a.bin_list[cb + gst[u] + (i - gof[u])];
          gl[i] = val1;
        }
        wave_lds_sync();
        int rk0 = 0, rk1 = 0;  // rank sort (indices are distinct)
        for (int u = 0; u < n; u++) {
          const int g = gl[u];
          rk0 += g < val0;
          rk1 += g < val1;
        }
        // the sorted list goes to the candidate array, which the scan below
        // filters in place (a candidate's slot never lies above its source)
        if (lane < n) cl[rk0] = val0;
        if (lane + DPM_WAVE < n) cl[rk1] = val1;
        wave_lds_sync();
        gathered = true;
        nsrc = n;
      }
    }
    int c0 = 0;
    // a chunk's source cell and statistics, loaded one chunk ahead (a gathered
    // source is read before the compaction below can overwrite it: the
    // compacted slots lie below the chunk being tested)
    int cj_nx = lane < nsrc ? (gathered ? cl[lane] : cb + lane) : ce;
    CellStat<R> st_nx{};
    if (cj_nx < ce && cj_nx != ci) st_nx = a.st_in[cj_nx];
    while (c0 < nsrc) {  // wave-uniform: candidates in batches of at most kWaveCand
      int ncl = 0;
      while (c0 < nsrc && ncl + DPM_WAVE <= kWaveCand) {
        const int cj = cj_nx;
        const CellStat<R> stj = st_nx;
        {
          const int k1 = c0 + DPM_WAVE + lane;
          cj_nx = k1 < nsrc ? (gathered ? cl[k1] : cb + k1) : ce;
          if (cj_nx < ce && cj_nx != ci) st_nx = a.st_in[cj_nx];
        }
        bool cand = false;
        if (cj < ce && cj != ci) {
          // a superset test: the nearest image by the rounded product with
          // 1/L (where that is ambiguous, near a half box length, both images
          // lie beyond any reach below 0.4995 L; longer reaches take every
          // cell); the pair's exact shift is taken again for a candidate
          V3<R> sh = zero;
          if (pbc)
            sh = v3(L * round((Ci.x - stj.cx) * invL), L * round((Ci.y - stj.cy) * invL),
                    L * round((Ci.z - stj.cz) * invL));
          const V3<R> d = {stj.cx + sh.x - Ci.x, stj.cy + sh.y - Ci.y, stj.cz + sh.z - Ci.z};
          R reach = R(0);
          if (do_contact) {
            reach = Ri + stj.r + contact_margin(stj.r, stj.em);
            if (pbc && R(2) * reach >= R(0.999f) * L) cand = true;
          }
          if (do_adh) {
            const R l0m = fmax(L0i, fmax(stj.l0, R(1e-12f)));
            const R r2 = Ri + stj.r + R(2) * l0m;
            reach = fmax(reach, r2);
            if (pbc && R(2) * r2 >= R(0.999f) * L) cand = true;  // minimum image per pair
          }
          const R rr = reach * R(1.0001f) + R(1e-6f);
          cand = cand || vdot(d, d) <= rr * rr;
        }
        const unsigned long long cmask = __ballot(cand);
        if (cand) cl[ncl + lanes_below(cmask)] = cj;  // ascending
        ncl += __popcll(cmask);
        c0 += DPM_WAVE;
      }
      n_cand += (unsigned)ncl;
      wave_lds_sync();
      for (int qc = 0; qc < ncl; qc++) {
        const int cj2 = cl[qc];
        const CellStat<R> stj = a.st_in[cj2];
        // cell j's positions stay in registers through the list pass (lane l
        // holds vertices l, l + 64, ..., as its chunks of the pass); every
        // load is issued before the first use
        R qx[kChunks], qy[kChunks], qz[kChunks];
        {
          const size_t bj = (size_t)cj2 * nv;
#pragma unroll
          for (int k = 0; k < kChunks; k++) {
            const int v = lane + k * DPM_WAVE;
            qx[k] = qy[k] = qz[k] = R(0);
            if (v < nv) {
              qx[k] = a.xin[bj + v];
              qy[k] = a.yin[bj + v];
              qz[k] = a.zin[bj + v];
            }
          }
        }
        V3<R> sh = zero;
        if (pbc)
          sh = v3(image_shift(Ci.x - stj.cx, L, invL), image_shift(Ci.y - stj.cy, L, invL),
                  image_shift(Ci.z - stj.cz, L, invL));
        const V3<R> Cj = {stj.cx + sh.x, stj.cy + sh.y, stj.cz + sh.z};
        const R mj = contact_margin(stj.r, stj.em);
        const R rj = (stj.r + mj) * R(1.0001f) + R(1e-6f);
        const R L0j = fmax(stj.l0, R(1e-12f));
        const R l0m = fmax(L0i, L0j);
        const bool all_images = pbc && R(2) * (Ri + stj.r + R(2) * l0m) >= R(0.999f) * L;
        // every |p - pj| component stays below 0.49 L, and no pair needs an
        // image, when the unshifted centres and radii keep it below 0.48 L
        // (then the shift is 0 too)
        const bool no_image = !pbc || (fabs(Ci.x - stj.cx) + (Ri + stj.r) < R(0.48f) * L &&
                                       fabs(Ci.y - stj.cy) + (Ri + stj.r) < R(0.48f) * L &&
                                       fabs(Ci.z - stj.cz) + (Ri + stj.r) < R(0.48f) * L);
        // vertices of cell j within reach of this cell, and of this cell
        // within reach of cell j (|q - Ci| < 2 l0 + Ri, and reversed)
        const R rk = (Ri + R(2) * l0m) * R(1.0001f) + R(1e-6f);
        const R r2 = (stj.r + R(2) * l0m) * R(1.0001f) + R(1e-6f);
        // Separation along the centre line n = (C_j - C_i) / D: every vertex p
        // of this cell has (p - C_i).n <= hi and every point x of cell j's
        // (shifted) hull has (C_j - x).n <= hj, so |x - p| >= D - hi - hj
        // = min (q + sh - C_i).n - max (p - C_j).n - D over the vertices.
        const V3<R> dC = vsub(Cj, Ci);
        const R D = sqrt(vdot(dC, dC));
        const V3<R> nC = D > R(0) ? vscale(dC, R(1) / D) : zero;
        R hp = R(-3.0e38f), hq = R(3.0e38f);
        const V3<R> Cis = vsub(Ci, sh);
        // One pass: the separation, whether an own vertex is within contact
        // reach, and the adhesion lists.  Cell j's vertices in reach are
        // compacted to the front of nx/ny/nz (ascending).
        bool any_need = false;
        int nnear = 0, nmine = 0;
#pragma unroll
        for (int k = 0; k < kChunks; k++) {
          const int k0v = k * DPM_WAVE;
          if (k0v >= nv) break;  // wave-uniform
          const int v = k0v + lane;
          bool need = false, near = false, mine = false;
          const V3<R> q = {qx[k], qy[k], qz[k]};
          if (v < nv) {
            const V3<R> P = lds_v(sx, sy, sz, v);
            const V3<R> dp = vsub(P, Cj);
            const V3<R> ds = vsub(q, Cis);
            hp = fmax(hp, vdot(dp, nC));
            hq = fmin(hq, vdot(ds, nC));
            const R d2 = vdot(dp, dp);
            need = do_contact && d2 <= rj * rj;
            if (do_adh) {
              mine = all_images || d2 <= r2 * r2;
              near = all_images || vdot(ds, ds) <= rk * rk;
            }
          }
          any_need = any_need || __ballot(need) != 0ull;
          const unsigned long long mn = __ballot(near), mm = __ballot(mine);
          if (near) {
            const int jn = nnear + lanes_below(mn);
            nx[jn] = q.x;
            ny[jn] = q.y;
            nz[jn] = q.z;
          }
          const int slot = nmine + lanes_below(mm);
          if (mine) il[slot] = (unsigned short)v;
          nnear += __popcll(mn);
          nmine += __popcll(mm);
        }
        hp = wave_max_all(hp);
        hq = wave_min_all(hq);
        // the separation less slack for the float sums (relative 1e-4 of the
        // cells' extent, 1e-6 of the coordinates); -1 when it is not defined
        const R Mc = ((fabs(Ci.x) + fabs(Ci.y)) + fabs(Ci.z)) + ((fabs(Cj.x) + fabs(Cj.y)) + fabs(Cj.z));
        const R gap = D > R(0) ? ((hq - hp) - D) - (R(1e-4f) * ((Ri + stj.r) + D) + R(1e-6f) * Mc) : R(-1);
        // Adhesion: no pair of vertices closer than 2 l0 when gap >= 2 l0 (on
        // pairs that take no image; the others are always evaluated).
        if (do_adh && nnear > 0 && nmine > 0 && (all_images || !no_image || !(gap >= R(2) * l0m))) {  // wave-uniform
          wave_lds_sync();
          // (own vertex, slice of near vertices) per lane; the slices are
          // summed in order afterwards
          const int S = nmine >= DPM_WAVE ? 1 : max(1, min(nnear, DPM_WAVE / nmine));
          const int per = (nnear + S - 1) / S;
          const R half_l = R(0.49f) * L;
          const AdhCut<R> cut_own(L0i * R(2)), cut_rev(L0j * R(2));
          const double kh = 0.5 * (double)Kat, inv_l0i = sti.inv_l0, inv_l0j = stj.inv_l0;
          const float kh_f = (float)kh, inv_l0i_f = (float)inv_l0i, inv_l0j_f = (float)inv_l0j;
          for (int tw = lane; tw < S * nmine; tw += DPM_WAVE) {
            const int iv = tw % nmine, sl = tw / nmine;
            const V3<R> Pv = lds_v(sx, sy, sz, (int)il[iv]);
            V3<R> Fa = zero;
            int cnt_own = 0;
            const int m1 = min(nnear, (sl + 1) * per);
            for (int m = sl * per; m < m1; m++) {
              const V3<R> pj = lds_v(nx, ny, nz, m);
              // delta = pj - p and p - pj, minimum image per component
              // (Cell3D_Kernel.cl:341-346); decisions in the reference's
              // arithmetic.  Without an image, p - pj is -(pj - p) exactly and
              // the two squared distances are the same number; a float step
              // then takes the accepted terms in float (adhesion_term_near),
              // pairs across the boundary in double on the same image.
              V3<R> dl = vsub(pj, Pv);
              V3<R> nl = zero, nr = zero;
              R d2r;
              bool imaged = false;
              if (!no_image && !(fabs(dl.x) < half_l && fabs(dl.y) < half_l && fabs(dl.z) < half_l)) {
                V3<R> dr = vsub(Pv, pj);
                nl = v3(round(dl.x / L), round(dl.y / L), round(dl.z / L));
                nr = v3(round(dr.x / L), round(dr.y / L), round(dr.z / L));
                dl = vsub(dl, v3(L * nl.x, L * nl.y, L * nl.z));
                dr = vsub(dr, v3(L * nr.x, L * nr.y, L * nr.z));
                d2r = vdot(dr, dr);
                imaged = true;
              } else {
                d2r = vdot(dl, dl);
              }
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
              if (cut_rev(d2r)) {
                Fa = vadd(Fa, adhesion_term(Pv, pj, nr, L, Kat, L0j, kh, inv_l0j));
                if (PROBE && a.p_adh_hash)
                  atomicAdd(&a.p_adh_hash[2 * (base + il[iv]) + 1], vertex_key(pj.x, pj.y, pj.z));
              }
            }
            ut[sl * nmine + iv] = Q4<R>{Fa.x, Fa.y, Fa.z, (R)cnt_own};
          }
          wave_lds_sync();
          for (int slot = lane; slot < nmine; slot += DPM_WAVE) {
            const int v = il[slot];
            V3<R> F = lds_v(fx, fy, fz, v);
            int cnt_own = 0;
            for (int s2 = 0; s2 < S; s2++) {
              const Q4<R> q4 = ut[s2 * nmine + slot];
              F = vadd(F, V3<R>{q4.x, q4.y, q4.z});
              cnt_own += (int)q4.w;
            }
            fx[v] = F.x;
            fy[v] = F.y;
            fz[v] = F.z;
            if (PROBE && a.p_adh) a.p_adh[base + v] += cnt_own;
          }
        }
        // Contact: a point at least m_j from every face of cell j has winding
        // number 0 (contact_margin), so the pass is needed only below that
        // separation; it stages cell j's whole mesh.
        if (do_contact && any_need && !(gap >= mj)) {  // wave-uniform
          wave_lds_sync();
          wave_stage_cell(nx, ny, nz, a.xin, a.yin, a.zin, (size_t)cj2 * nv, nv, lane);
          wave_lds_sync();
          #pragma unroll 1
          for (int k0v = 0; k0v < nv; k0v += DPM_WAVE) {
            const int v = k0v + lane;
            bool need = false;
            V3<R> P = zero;
            if (v < nv) {
              P = lds_v(sx, sy, sz, v);
              const V3<R> d = vsub(P, Cj);
              need = vdot(d, d) <= rj * rj;
            }
            unsigned long long nm = __ballot(need);
            R wk = R(0);
            while (nm) {  // wave-uniform: one point at a time
              const int bb = __ffsll((long long)nm) - 1;
              nm &= nm - 1ull;
              const V3<R> Pb = lds_v(sx, sy, sz, k0v + bb);
              // a point beyond the margin from j's hull has winding number 0
              if (outside_by_margin(nx, ny, nz, nv, stj.cx, stj.cy, stj.cz, Cj, vsub(Pb, Cj), mj, stj.r)) continue;
              n_wind++;
              const R w = winding_w(nx, ny, nz, sf, ut, nv, nf, sh, Pb);
              if (lane == bb) wk = w;
            }
            if (need) {
              const size_t pidx = (base + v) * (size_t)a.ncmax + (size_t)(cj2 - cb);
              if (PROBE && a.p_winding) a.p_winding[pidx] = (float)wk;
              if (!(fabs(wk) < R(1e-6f))) {
                const V3<R> dir = vnormalize(vsub(Ci, P));
                const R mag = fabs(wk) * R(0.5f) * Kre;
                const V3<R> F = vadd(lds_v(fx, fy, fz, v), vscale(dir, mag));
                fx[v] = F.x;
                fy[v] = F.y;
                fz[v] = F.z;
                n_contact++;
                if (a.w_out) a.w_out[base + v] += fabs(wk);
                if (a.rec_cell) {  // the first free of the vertex's partner slots
                  const size_t r0 = (base + v) * DPM3D_MAXPART;
                  int q = 0;
                  while (q < DPM3D_MAXPART && a.rec_cell[r0 + q] >= 0) q++;
                  if (q < DPM3D_MAXPART) {
                    const size_t r = r0 + q;
                    a.rec_cell[r] = cj2;
                    a.rec_w[r] = wk;
                    a.rec_shift[3 * r] = sh.x;
                    a.rec_shift[3 * r + 1] = sh.y;
                    a.rec_shift[3 * r + 2] = sh.z;
                  } else {
                    atomicOr(a.rec_overflow, 1);
                  }
                }
                if (PROBE && a.p_contact) a.p_contact[pidx] = 1;
                if (PROBE && a.p_contact_key) {  // count, then key sum of the partner cells
                  atomicAdd(&a.p_contact_key[2 * (base + v)], 1u);
                  atomicAdd(&a.p_contact_key[2 * (base + v) + 1], cell_key(cj2 - cb));
                }
              }
            }
          }
        }
        wave_lds_sync();  // the candidate arrays are rewritten next
      }
    }
  }

  // ------------------------------------------------------- Euler and stats
  double px = 0.0, py = 0.0, pz = 0.0;
  #pragma unroll 1
  for (int v = lane; v < nv; v += DPM_WAVE) {
    const V3<R> F = lds_v(fx, fy, fz, v);
    if (a.flags & F_WRITE_FORCES) {
      a.fx[base + v] = F.x;
      a.fy[base + v] = F.y;
      a.fz[base + v] = F.z;
    }
    if (a.flags & F_MOVE) {
      const V3<R> P = lds_v(sx, sy, sz, v);
      V3<R> Pn;
      if (a.xlo) {
        // x + lo + F dt, the displacement rounded once onto the residual
        R ex, ey, ez;
        Pn = v3(two_sum(P.x, fma(F.x, a.dt, a.xlo[base + v]), ex),
                two_sum(P.y, fma(F.y, a.dt, a.ylo[base + v]), ey),
                two_sum(P.z, fma(F.z, a.dt, a.zlo[base + v]), ez));
        a.xlo[base + v] = ex;
        a.ylo[base + v] = ey;
        a.zlo[base + v] = ez;
      } else {
        Pn = v3(fma(F.x, a.dt, P.x), fma(F.y, a.dt, P.y), fma(F.z, a.dt, P.z));
      }
      a.xout[base + v] = Pn.x;
      a.yout[base + v] = Pn.y;
      a.zout[base + v] = Pn.z;
      sx[v] = Pn.x;
      sy[v] = Pn.y;
      sz[v] = Pn.z;
      px += (double)Pn.x;
      py += (double)Pn.y;
      pz += (double)Pn.z;
    }
  }
  if (a.flags & F_MOVE) {  // wave-uniform: statistics for the next step (cell_stats)
    wave_lds_sync();
    px = wsum_all(px);
    py = wsum_all(py);
    pz = wsum_all(pz);
    const V3<R> C = {(R)(px / nv), (R)(py / nv), (R)(pz / nv)};
    R r2 = R(0);  // max |d|^2, then one square root (monotonic, correctly rounded)
    #pragma unroll 1
    for (int v = lane; v < nv; v += DPM_WAVE) {
      const V3<R> d = vsub(lds_v(sx, sy, sz, v), C);
      r2 = fmax(r2, vdot(d, d));
    }
    const R r = sqrt(r2);
    double vol = 0.0;
    R e2 = R(0);
    #pragma unroll 1
    for (int f = lane; f < nf; f += DPM_WAVE) {
      int i0, i1, i2;
      wface(sf, f, i0, i1, i2);
      const V3<R> P0 = vsub(lds_v(sx, sy, sz, i0), C);
      const V3<R> P1 = vsub(lds_v(sx, sy, sz, i1), C);
      const V3<R> P2 = vsub(lds_v(sx, sy, sz, i2), C);
      vol += (double)vdot(vcross(P0, P1), P2);
      const V3<R> e0 = vsub(P1, P0), e1 = vsub(P2, P1), ee = vsub(P0, P2);
      e2 = fmax(e2, fmax(vdot(e0, e0), fmax(vdot(e1, e1), vdot(ee, ee))));
    }
    const R rd = wave_max_all(r), e2d = wave_max_all(e2);  // maxima are exact in either precision
    vol = wsum_all(vol);
    if (lane == 0) {
      const R em = sqrt((R)e2d) * R(1.000001f);
      a.st_out[ci] = make_cell_stat<R>(C.x, C.y, C.z, (R)rd * R(1.0000005f) + R(1e-30f), em, l0i);
      a.vol_out[ci] = fabs(vol / 6.0);
      a.em_out[ci] = em;
    }
  }
  if (a.counters) {
    // n_contact counts this lane's contacts; n_wind and n_cand are wave totals
    const unsigned long long c = (unsigned long long)wsum_all((double)n_contact);
    if (lane == 0) {
      unsigned long long* slot = a.counters + 4 * ((blockIdx.x * cpw + wv) % DPM_COUNTER_SLOTS);
      if (ci == 0) atomicAdd(&slot[0], 1ull);
      if (n_wind) atomicAdd(&slot[1], (unsigned long long)n_wind);
      if (c) atomicAdd(&slot[2], c);
      if (n_cand) atomicAdd(&slot[3], (unsigned long long)n_cand);
    }
  }
}

extern "C" __global__ void __launch_bounds__(256, 4) k3d_wstep_f32(const Step3D<float> a) {
  k3d_wstep_body<float, 0, 0>(a);
}
// the icosphere of subdivision 2 (162 vertices, 320 faces); the production
// instance has the diagnostic outputs compiled out (their registers would
// spill), the _probe instance of the same source records them
extern "C" __global__ void __launch_bounds__(256, 4) k3d_wstep_f32_v162(const Step3D<float> a) {
  k3d_wstep_body<float, 162, 320, false>(a);
}
extern "C" __global__ void __launch_bounds__(256, 4) k3d_wstep_f32_v162_probe(const Step3D<float> a) {
  k3d_wstep_body<float, 162, 320, true>(a);
}
extern "C" __global__ void __launch_bounds__(256) k3d_wstep_f64(const Step3D<double> a) { k3d_wstep_body<double, 0, 0>(a); }
