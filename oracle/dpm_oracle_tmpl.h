/*
 * TEST INFRASTRUCTURE ONLY — the oracle.  Nothing in the product links, loads
 * or calls this code; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it, and only as the checker.
 *
 * Body of the CPU restatement of the reference DPM kernels, written once and
 * instantiated twice by dpm_oracle.c:
 *   R = float   -> "ref32": the reference arithmetic.  Variables are float and
 *                  every unsuffixed literal of the reference is kept as a
 *                  double literal here, so C's usual arithmetic conversions
 *                  (FLT_EVAL_METHOD 0 on x86-64 SSE) reproduce the same
 *                  float/double promotions the OpenCL C kernels get.
 *   R = double -> "ref64": the same formulas, all in double.
 * Built with -ffp-contract=off: every a*b+c is two rounded operations.
 *
 * Stated choices where the reference leaves arithmetic to its OpenCL runtime
 * (parity unpinned at the bit level; pinned by the tests to closed forms,
 * invariants and finite differences instead):
 *   - normalize(v) = v * (1 / sqrt(dot(v,v))), and normalize(0) = 0;
 *   - dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z, no fused multiply-add;
 *   - length(v) = sqrt(dot(v,v));
 *   - atan2 = the C library's atan2f / atan2;
 *   - Euler update x += F*dt is fused (fmaf), as clang emits fmuladd there;
 *   - the 3D volume is the volume of the current positions (the reference's
 *     per-cell volume buffer is written by one work-item and read by the
 *     others without a global barrier, Cell3D_Kernel.cl:78-83; the intended
 *     value is the current volume).
 */

#ifndef FN
#error "include from dpm_oracle.c"
#endif

/* ----------------------------------------------------------------- helpers */

static inline R FN(dot2)(R ax, R ay, R bx, R by) { return ax * bx + ay * by; }
static inline R FN(dot3)(const R a[3], const R b[3]) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
static inline void FN(cross3)(const R a[3], const R b[3], R c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static inline void FN(normalize3)(const R a[3], R u[3]) {
  R d = FN(dot3)(a, a);
  if (d == (R)0) { u[0] = u[1] = u[2] = (R)0; return; }
  R inv = (R)1 / SQRT(d);
  u[0] = a[0] * inv; u[1] = a[1] * inv; u[2] = a[2] * inv;
}
static inline void FN(normalize2)(R x, R y, R* ux, R* uy) {
  R d = FN(dot2)(x, y, x, y);
  if (d == (R)0) { *ux = *uy = (R)0; return; }
  R inv = (R)1 / SQRT(d);
  *ux = x * inv; *uy = y * inv;
}

/* Pair-set fingerprints (diagnostics, no reference counterpart): a partner
 * vertex is identified by the bit patterns of its coordinates,
 * key = f(x ^ f(y ^ f(z))) with f the 32-bit finaliser of MurmurHash3 (a
 * bijection of 32-bit words; a double contributes f(low word) ^ high word),
 * and a vertex's partner set by the wrapping 32-bit sum of its partners'
 * keys, which does not depend on the order the pairs are found in. */
static inline unsigned FN(fmix32)(unsigned h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
static inline unsigned FN(rbits)(R v) {
  if (sizeof(R) == 4) {
    unsigned b;
    memcpy(&b, &v, 4);
    return b;
  }
  unsigned long long b = 0;
  memcpy(&b, &v, sizeof(R));
  return FN(fmix32)((unsigned)b) ^ (unsigned)(b >> 32);
}
static inline unsigned FN(vkey2)(R x, R y) {
  return FN(fmix32)(FN(rbits)(x) ^ FN(fmix32)(FN(rbits)(y)));
}
static inline unsigned FN(vkey3)(R x, R y, R z) {
  return FN(fmix32)(FN(rbits)(x) ^ FN(fmix32)(FN(rbits)(y) ^ FN(fmix32)(FN(rbits)(z))));
}

/* ======================================================================= 2D
 * Layout: padded AoS like the reference, V[(ci*stride + vi)*2 + {0,1}]
 * (Tissue2D.cpp:117-140).  Only vi < NV[ci] is read or written.
 */

/* GetCOM, Cell2D_kernel.cl:3-11: float2 sum then divide by NV. */
static void FN(com2)(const R* V, int stride, int ci, int nv, R* cx, R* cy) {
  R sx = 0, sy = 0;
  for (int i = 0; i < nv; i++) {
    sx += V[(ci * stride + i) * 2 + 0];
    sy += V[(ci * stride + i) * 2 + 1];
  }
  *cx = sx / (R)nv;
  *cy = sy / (R)nv;
}

/* Point-in-polygon crossing parity of vertex p against cell cj with the
 * reference's periodic quirk (floor in x, round in y, only when a component
 * exceeds L), Cell2D_kernel.cl:171-196. */
static int FN(crossing_parity)(const R* V, int stride, const int* NV, int cj,
                               R px, R py, int PBC, R L) {
  int odd = 0;
  int n = NV[cj];
  for (int i = 0, j = n - 1; i < n; j = i++) {
    R dix = px - V[(cj * stride + i) * 2 + 0];
    R diy = py - V[(cj * stride + i) * 2 + 1];
    R djx = px - V[(cj * stride + j) * 2 + 0];
    R djy = py - V[(cj * stride + j) * 2 + 1];
    if (PBC) {
      if (FABS(dix) > L || FABS(djx) > L) {
        dix -= L * FLOOR(dix / L);
        djx -= L * FLOOR(djx / L);
      }
      if (FABS(diy) > L || FABS(djy) > L) {
        diy -= L * ROUND(diy / L);
        djy -= L * ROUND(djy / L);
      }
    }
    if ((diy > 0) != (djy > 0)) {
      R t = (djx - dix) * (0 - diy) / (djy - diy) + dix;
      if (0 < t) odd = !odd;
    }
  }
  return odd;
}

/* One evaluation of all 2D force terms at the given positions, in the
 * reference's kernel order Area, Perimeter, Bending, Attraction, Repulsion
 * (Tissue2D.cpp:216-222).  F is overwritten.  overlap (optional, per padded
 * slot) receives 1 where the repulsion pair (vertex, some other cell) exists;
 * attr_count (optional) the number of attraction partners of each vertex. */
int FN(ora2d_forces)(int ncells, int stride, const int* NV, const R* V,
                     const R* Ka, const R* Kl, const R* Kb, const R* a0,
                     const R* l0, const R* r0, R L, int PBC, R Kre, R Kat,
                     unsigned mask, R* F, int* overlap, int* attr_count,
                     unsigned* attr_hash) {
  for (int k = 0; k < ncells * stride * 2; k++) F[k] = 0;
  if (overlap)
    for (int k = 0; k < ncells * stride; k++) overlap[k] = 0;
  if (attr_count)
    for (int k = 0; k < ncells * stride; k++) attr_count[k] = 0;
  if (attr_hash)
    for (int k = 0; k < ncells * stride; k++) attr_hash[k] = 0;

  ORA_PARALLEL_FOR
  for (int ci = 0; ci < ncells; ci++) {
    const int n = NV[ci];
    const int base = ci * stride;
#define X(k) V[(k) * 2 + 0]
#define Y(k) V[(k) * 2 + 1]
    /* AreaForceUpdates, Cell2D_kernel.cl:13-42.  The area is recomputed per
     * work-item in the reference; it is the same value for every vertex. */
    if (mask & 1u) {
      R Area = 0.0f;
      for (int vj = 0; vj < n; vj++) {
        int idx = base + vj;
        int im = (vj == 0) ? (base + n - 1) : (idx - 1);
        Area += 0.5 * ((X(im) + X(idx)) * (Y(im) - Y(idx)));
      }
      if (Area < 0.0) Area = -Area;
      R areaStrain = (Area / a0[ci]) - 1.0;
      for (int vi = 0; vi < n; vi++) {
        int index = base + vi;
        int im1 = (vi == 0) ? (base + n - 1) : (index - 1);
        int ip1 = (vi == n - 1) ? base : (index + 1);
        F[index * 2 + 0] += (Ka[ci] / SQRT(a0[ci])) * 0.5 * areaStrain * (Y(im1) - Y(ip1));
        F[index * 2 + 1] += (Ka[ci] / SQRT(a0[ci])) * 0.5 * areaStrain * (X(im1) - X(ip1));
      }
    }
    /* Gradient area force (no reference counterpart): -dE/dx of
     * E = Ka sqrt(a0)/2 (|A|/a0 - 1)^2 for either orientation; equals the
     * reference's x component on counter-clockwise cells and flips its y
     * component. */
    if (mask & 32u) {
      double As = 0.0;
      for (int vj = 0; vj < n; vj++) {
        int idx = base + vj;
        int im = (vj == 0) ? (base + n - 1) : (idx - 1);
        As += 0.5 * ((double)(X(im) + X(idx)) * (double)(Y(im) - Y(idx)));
      }
      const R sgn = As < 0.0 ? (R)-1 : (R)1;
      const R s = (R)(fabs(As) / (double)a0[ci] - 1.0);
      const R c = Ka[ci] / SQRT(a0[ci]);
      for (int vi = 0; vi < n; vi++) {
        int index = base + vi;
        int im1 = (vi == 0) ? (base + n - 1) : (index - 1);
        int ip1 = (vi == n - 1) ? base : (index + 1);
        F[index * 2 + 0] += -c * s * sgn * (R)0.5 * (Y(im1) - Y(ip1));
        F[index * 2 + 1] += -c * s * sgn * (R)0.5 * (X(ip1) - X(im1));
      }
    }
    for (int vi = 0; vi < n; vi++) {
      int index = base + vi;
      int im1 = (vi == 0) ? (base + n - 1) : (index - 1);
      int ip1 = (vi == n - 1) ? base : (index + 1);
      /* PerimeterForceUpdates, Cell2D_kernel.cl:92-119. */
      if (mask & 2u) {
        R lvx = X(ip1) - X(index), lvy = Y(ip1) - Y(index);
        R lvmx = X(index) - X(im1), lvmy = Y(index) - Y(im1);
        R length = SQRT(FN(dot2)(lvx, lvy, lvx, lvy));
        R lengthm = SQRT(FN(dot2)(lvmx, lvmy, lvmx, lvmy));
        R ulvx = lvx / length, ulvy = lvy / length;
        R ulvmx = lvmx / lengthm, ulvmy = lvmy / lengthm;
        R dli = length / l0[ci] - 1.0;
        R dlim1 = lengthm / l0[ci] - 1.0;
        R pref = Kl[ci] * SQRT(a0[ci] / l0[ci]);
        F[index * 2 + 0] += pref * (dli * ulvx - dlim1 * ulvmx);
        F[index * 2 + 1] += pref * (dli * ulvy - dlim1 * ulvmy);
      }
      /* BendingForceUpdates, Cell2D_kernel.cl:44-90. */
      if (mask & 4u) {
        int ip2 = (vi == n - 1) ? base + 1 : ((vi == n - 2) ? base : index + 2);
        int im2 = (vi == 0) ? (base + n - 2) : ((vi == 1) ? (base + n - 1) : index - 2);
        R lvx = X(ip1) - X(index), lvy = Y(ip1) - Y(index);
        R lvxm = X(index) - X(im1), lvym = Y(index) - Y(im1);
        R six = lvx - lvxm, siy = lvy - lvym;
        R sixp = (X(ip2) - X(ip1)) - lvx, siyp = (Y(ip2) - Y(ip1)) - lvy;
        R sixm = lvxm - (X(im1) - X(im2)), siym = lvym - (Y(im1) - Y(im2));
        F[index * 2 + 0] += Kb[ci] * (2.0 * six - sixm - sixp);
        F[index * 2 + 1] += Kb[ci] * (2.0 * siy - siym - siyp);
      }
      /* AttractionForceUpdate, Cell2D_kernel.cl:222-268. */
      if (mask & 8u) {
        for (int cj = 0; cj < ncells; cj++) {
          if (cj == ci) continue;
          for (int vj = 0; vj < NV[cj]; vj++) {
            int j = cj * stride + vj;
            R rx = X(j) - X(index), ry = Y(j) - Y(index);
            if (PBC) {
              rx -= L * ROUND(rx / L);
              ry -= L * ROUND(ry / L);
            }
            R dist = SQRT(FN(dot2)(rx, ry, rx, ry));
            if (dist < l0[ci]) {
              R ftmp = Kat / (R)NV[ci] * dist / l0[ci];
              R ux, uy;
              FN(normalize2)(rx, ry, &ux, &uy);
              F[index * 2 + 0] += ftmp * ux;
              F[index * 2 + 1] += ftmp * uy;
              if (attr_count) attr_count[index]++;
              if (attr_hash) attr_hash[index] += FN(vkey2)(X(j), Y(j));
            }
          }
        }
      }
      /* Vertex soft-disk repulsion (no reference counterpart): vertices of
       * different cells closer than sigma = (l0_i + l0_j)/2 repel with
       * U = Kre/2 (1 - d/sigma)^2, minimum image as the attraction. */
      if (mask & 64u) {
        for (int cj = 0; cj < ncells; cj++) {
          if (cj == ci) continue;
          const R sig = (R)0.5 * (l0[ci] + l0[cj]);
          for (int vj = 0; vj < NV[cj]; vj++) {
            int j = cj * stride + vj;
            R rx = X(j) - X(index), ry = Y(j) - Y(index);
            if (PBC) {
              rx -= L * ROUND(rx / L);
              ry -= L * ROUND(ry / L);
            }
            R dist = SQRT(FN(dot2)(rx, ry, rx, ry));
            if (dist < sig && dist > 0) {
              R ftmp = Kre / sig * (1 - dist / sig) / dist;
              F[index * 2 + 0] -= ftmp * rx;
              F[index * 2 + 1] -= ftmp * ry;
            }
          }
        }
      }
      /* RepulsionForceUpdate, Cell2D_kernel.cl:121-220: overlap with the
       * first cell whose crossing parity is odd, then a push toward the own
       * centre of mass. */
      if (mask & 16u) {
        int overlaps = 0;
        for (int cj = 0; cj < ncells && !overlaps; cj++) {
          if (cj == ci) continue;
          overlaps = FN(crossing_parity)(V, stride, NV, cj, X(index), Y(index), PBC, L);
        }
        if (overlaps) {
          if (overlap) overlap[index] = 1;
          R cx, cy;
          FN(com2)(V, stride, ci, n, &cx, &cy);
          R dx = cx - X(index), dy = cy - Y(index);
          if (PBC) {
            dx -= L * ROUND(dx / L);
            dy -= L * ROUND(dy / L);
          }
          R dist = SQRT(FN(dot2)(dx, dy, dx, dy));
          R xij = dist / (2 * r0[ci]);
          R ftmp = Kre * (1 - xij);
          R ux, uy;
          FN(normalize2)(dx, dy, &ux, &uy);
          F[index * 2 + 0] += 0.5f * ftmp * ux;
          F[index * 2 + 1] += 0.5f * ftmp * uy;
        }
      }
    }
#undef X
#undef Y
  }
  return 0;
}

/* Tissue2D::CLEulerUpdate's step loop, Tissue2D.cpp:215-229 with
 * EulerUpdate, Cell2D_kernel.cl:270-281.  Fout = forces of the last step
 * (read before the Euler kernel zeroes them). */
int FN(ora2d_run)(int ncells, int stride, const int* NV, R* V, const R* Ka,
                  const R* Kl, const R* Kb, const R* a0, const R* l0,
                  const R* r0, R L, int PBC, R Kre, R Kat, unsigned mask,
                  int nsteps, R dt, R* Fout) {
  R* F = (R*)calloc((size_t)ncells * stride * 2, sizeof(R));
  if (!F) return -1;
  for (int s = 0; s < nsteps; s++) {
    FN(ora2d_forces)(ncells, stride, NV, V, Ka, Kl, Kb, a0, l0, r0, L, PBC, Kre,
                     Kat, mask, F, NULL, NULL, NULL);
    for (int ci = 0; ci < ncells; ci++)
      for (int vi = 0; vi < NV[ci]; vi++) {
        int k = (ci * stride + vi) * 2;
        V[k + 0] = FMA(F[k + 0], dt, V[k + 0]);
        V[k + 1] = FMA(F[k + 1], dt, V[k + 1]);
      }
  }
  if (Fout) memcpy(Fout, F, sizeof(R) * (size_t)ncells * stride * 2);
  free(F);
  return 0;
}

/* Energy of the conservative 2D terms (no reference counterpart), in
 * double: area (mask 1 or 32) Ka sqrt(a0)/2 (|A|/a0 - 1)^2, perimeter (2)
 * sum_e Kl sqrt(a0/l0) l0/2 (l/l0 - 1)^2, bending (4) Kb/2 sum |s_i|^2,
 * vertex repulsion (64) sum over pairs Kre/2 (1 - d/sigma)^2.  out[0] = total,
 * out[1] = the vertex-repulsion part. */
void FN(ora2d_energy)(int ncells, int stride, const int* NV, const R* V,
                      const R* Ka, const R* Kl, const R* Kb, const R* a0,
                      const R* l0, R L, int PBC, R Kre, unsigned mask,
                      double* out) {
  double U = 0.0, Ui = 0.0;
#define X(k) ((double)V[(k) * 2 + 0])
#define Y(k) ((double)V[(k) * 2 + 1])
  for (int ci = 0; ci < ncells; ci++) {
    const int n = NV[ci], base = ci * stride;
    if (mask & 33u) {
      double As = 0.0;
      for (int vj = 0; vj < n; vj++) {
        int idx = base + vj, im = (vj == 0) ? (base + n - 1) : (idx - 1);
        As += 0.5 * ((X(im) + X(idx)) * (Y(im) - Y(idx)));
      }
      const double s = fabs(As) / (double)a0[ci] - 1.0;
      U += 0.5 * (double)Ka[ci] * sqrt((double)a0[ci]) * s * s;
    }
    for (int vi = 0; vi < n; vi++) {
      const int k = base + vi;
      const int kp = (vi == n - 1) ? base : k + 1, km = (vi == 0) ? base + n - 1 : k - 1;
      if (mask & 2u) {
        const double dx = X(kp) - X(k), dy = Y(kp) - Y(k);
        const double e = sqrt(dx * dx + dy * dy) / (double)l0[ci] - 1.0;
        U += 0.5 * (double)Kl[ci] * sqrt((double)a0[ci] / (double)l0[ci]) * (double)l0[ci] * e * e;
      }
      if (mask & 4u) {
        const double sx = (X(kp) - X(k)) - (X(k) - X(km)), sy = (Y(kp) - Y(k)) - (Y(k) - Y(km));
        U += 0.5 * (double)Kb[ci] * (sx * sx + sy * sy);
      }
      if (mask & 64u)
        for (int cj = ci + 1; cj < ncells; cj++) {
          const double sig = 0.5 * ((double)l0[ci] + (double)l0[cj]);
          for (int vj = 0; vj < NV[cj]; vj++) {
            const int j = cj * stride + vj;
            double rx = X(j) - X(k), ry = Y(j) - Y(k);
            if (PBC) {
              rx -= (double)L * round(rx / (double)L);
              ry -= (double)L * round(ry / (double)L);
            }
            const double d = sqrt(rx * rx + ry * ry);
            if (d < sig) Ui += 0.5 * (double)Kre * (1.0 - d / sig) * (1.0 - d / sig);
          }
        }
    }
  }
#undef X
#undef Y
  out[0] = U + Ui;
  out[1] = Ui;
}

/* FIRE minimisation of one 2D tissue with the reference force law, no
 * reference counterpart (the reference only has the Euler step).  Bitzek,
 * Koskinen, Gaehler, Moseler & Gumbsch, PRL 97, 170201 (2006), with the
 * semi-implicit Euler MD step (unit masses), iteration by iteration:
 *   F = F(x); with u_clear >= 0 stop ("cleared") if the vertex-repulsion
 *   energy per cell is at most u_clear; stop if max_v |F_v| < ftol;
 *   P = sum F.v, |F|^2, |v|^2 (sums in double);
 *   P > 0: v = (1 - a) v + a |v| F/|F|; if n_pos > n_min: dt = min(dt f_inc,
 *          dt_max), a = a f_a; n_pos++;
 *   P < 0: v = 0, dt = dt f_dec, a = a0, n_pos = 0;
 *   (P = 0 happens only with v = 0 or F = 0, e.g. at the start, and changes
 *    nothing;)
 *   v += F dt; x += v dt.
 * V: padded positions (updated), Vel: padded velocities (in/out).
 * fp = {dt0, dt_max, ftol, f_inc, f_dec, alpha0, f_alpha, u_clear}; ip =
 * {max_iter, n_min}.  out = {iterations, converged, fmax, dt, alpha, n_pos,
 * cleared}. */
int FN(ora2d_fire)(int ncells, int stride, const int* NV, R* V, R* Vel,
                   const R* Ka, const R* Kl, const R* Kb, const R* a0,
                   const R* l0, const R* r0, R L, int PBC, R Kre, R Kat,
                   unsigned mask, const double* fp, const int* ip, double* out) {
  const size_t n2 = (size_t)ncells * stride * 2;
  R* F = (R*)calloc(n2, sizeof(R));
  if (!F) return -1;
  double dt = fp[0], alpha = fp[5], fmax = 0.0;
  int npos = 0, it = 0, conv = 0, cleared = 0;
  for (it = 0; it < ip[0]; it++) {
    FN(ora2d_forces)(ncells, stride, NV, V, Ka, Kl, Kb, a0, l0, r0, L, PBC, Kre,
                     Kat, mask, F, NULL, NULL, NULL);
    double P = 0.0, F2 = 0.0, v2 = 0.0;
    fmax = 0.0;
    for (int ci = 0; ci < ncells; ci++)
      for (int vi = 0; vi < NV[ci]; vi++) {
        const size_t k = ((size_t)ci * stride + vi) * 2;
        const double fx = F[k], fy = F[k + 1], vx = Vel[k], vy = Vel[k + 1];
        const double fn = sqrt(fx * fx + fy * fy);
        if (fn > fmax) fmax = fn;
        P += fx * vx + fy * vy;
        F2 += fx * fx + fy * fy;
        v2 += vx * vx + vy * vy;
      }
    if (fp[7] >= 0.0) {
      double eo[2];
      FN(ora2d_energy)(ncells, stride, NV, V, Ka, Kl, Kb, a0, l0, L, PBC, Kre, 64u, eo);
      if (eo[1] / ncells <= fp[7]) { cleared = 1; break; }
    }
    if (fmax < fp[2]) { conv = 1; break; }
    double mix_v = 1.0, mix_f = 0.0;
    int zero_v = 0;
    if (P > 0.0) {
      mix_v = 1.0 - alpha;
      mix_f = alpha * sqrt(v2 / F2);
      if (npos > ip[1]) {
        dt = fmin(dt * fp[3], fp[1]);
        alpha *= fp[6];
      }
      npos++;
    } else if (P < 0.0) {
      zero_v = 1;
      dt *= fp[4];
      alpha = fp[5];
      npos = 0;
    }
    for (int ci = 0; ci < ncells; ci++)
      for (int vi = 0; vi < NV[ci]; vi++) {
        const size_t k = ((size_t)ci * stride + vi) * 2;
        for (int c = 0; c < 2; c++) {
          R v = zero_v ? (R)0 : (R)(mix_v * (double)Vel[k + c] + mix_f * (double)F[k + c]);
          v = (R)((double)v + (double)F[k + c] * dt);
          Vel[k + c] = v;
          V[k + c] = (R)((double)V[k + c] + (double)v * dt);
        }
      }
  }
  out[0] = it; out[1] = conv; out[2] = fmax; out[3] = dt; out[4] = alpha; out[5] = npos;
  out[6] = cleared;
  free(F);
  return 0;
}

/* Jamming protocol (no reference counterpart): O'Hern, Silbert, Liu &
 * Nagel, PRE 68, 011306 (2003), as compression and bisection on
 * phi = sum(a0)/L^2.  Cells move rigidly so that their centres of mass
 * (double mean of the vertices) scale with the box; at each phi a FIRE
 * minimisation (ora2d_fire, velocities from rest) that ends early as
 * "cleared" once the vertex-repulsion energy per cell is at most u_jam;
 * jammed iff it was not cleared (converged to ftol or stopped at max_iter
 * with energy above u_jam).  phi grows by dphi from
 * phi0 until the first jammed state, then bisects between the largest
 * unjammed and the smallest jammed phi, restarting from the saved unjammed
 * state, until phi_hi - phi_lo < phi_tol or max_stages minimisations; ends
 * on the largest unjammed state.
 * jp = {phi0, dphi, phi_tol, u_jam}; ip = {max_iter, n_min, max_stages};
 * out = {phi_lo, phi_hi, stages, iterations, L, status, unconverged};
 * status 0 bracket within phi_tol, 1 phi0 jammed, 3 stopped at max_stages. */
static void FN(rescale2d)(int ncells, int stride, const int* NV, R* V, R* L, R Lnew) {
  const R s = Lnew / *L - (R)1;
  for (int ci = 0; ci < ncells; ci++) {
    double sx = 0.0, sy = 0.0;
    for (int vi = 0; vi < NV[ci]; vi++) {
      sx += (double)V[((size_t)ci * stride + vi) * 2];
      sy += (double)V[((size_t)ci * stride + vi) * 2 + 1];
    }
    const R cx = (R)(sx / NV[ci]), cy = (R)(sy / NV[ci]);
    for (int vi = 0; vi < NV[ci]; vi++) {
      V[((size_t)ci * stride + vi) * 2] += s * cx;
      V[((size_t)ci * stride + vi) * 2 + 1] += s * cy;
    }
  }
  *L = Lnew;
}

int FN(ora2d_jam)(int ncells, int stride, const int* NV, R* V, const R* Ka,
                  const R* Kl, const R* Kb, const R* a0, const R* l0,
                  const R* r0, R* Lio, int PBC, R Kre, R Kat, unsigned mask,
                  const double* fp, const int* ip, const double* jp, double* out) {
  const size_t n2 = (size_t)ncells * stride * 2;
  R* Vel = (R*)calloc(n2, sizeof(R));
  R* Save = (R*)malloc(n2 * sizeof(R));
  if (!Vel || !Save) return -1;
  double A0 = 0.0;
  for (int c = 0; c < ncells; c++) A0 += (double)a0[c];
  R L = *Lio;
  FN(rescale2d)(ncells, stride, NV, V, &L, (R)sqrt(A0 / jp[0]));
  double phi = jp[0], phi_lo = 0.0, phi_hi = 0.0, iters = 0.0, unconv = 0.0, status = 0.0;
  int stage = 0;
  const int fip[2] = {ip[0], ip[1]};
  double ffp[8];
  for (int k = 0; k < 7; k++) ffp[k] = fp[k];
  ffp[7] = jp[3];
  while (stage < ip[2]) {
    double fo[7];
    for (size_t k = 0; k < n2; k++) Vel[k] = 0;
    FN(ora2d_fire)(ncells, stride, NV, V, Vel, Ka, Kl, Kb, a0, l0, r0, L, PBC, Kre, Kat, mask, ffp, fip, fo);
    iters += fo[0];
    unconv += fo[1] == 0.0 && fo[6] == 0.0;
    stage++;
    const int jammed = fo[6] == 0.0;
    if (!jammed) {
      phi_lo = phi;
      memcpy(Save, V, n2 * sizeof(R));
    } else {
      phi_hi = phi;
      if (phi_lo == 0.0) { status = 1; break; }
      memcpy(V, Save, n2 * sizeof(R));
      L = (R)sqrt(A0 / phi_lo);
    }
    if (phi_hi > 0.0 && phi_hi - phi_lo < jp[2]) break;
    const double next = phi_hi > 0.0 ? 0.5 * (phi_lo + phi_hi) : phi + jp[1];
    FN(rescale2d)(ncells, stride, NV, V, &L, (R)sqrt(A0 / next));
    phi = next;
  }
  if (status == 0.0 && !(phi_hi > 0.0 && phi_hi - phi_lo < jp[2])) status = 3;  /* stage limit */
  if (phi_lo > 0.0 && phi != phi_lo) {
    memcpy(V, Save, n2 * sizeof(R));
    L = (R)sqrt(A0 / phi_lo);
  }
  *Lio = L;
  out[0] = phi_lo; out[1] = phi_hi; out[2] = stage; out[3] = iters; out[4] = (double)L;
  out[5] = status; out[6] = unconv;
  free(Vel);
  free(Save);
  return 0;
}

/* ======================================================================= 3D
 * Layout: AoS X[(ci*nv + v)*3 + {0,1,2}], faces[f*3 + k] shared by all cells
 * (Tissue3D.cpp:139-155).
 */

/* GetCOM, Cell3D_Kernel.cl:35-44: sum then multiply by 1/NV. */
static void FN(com3)(const R* X, int nv, int ci, R c[3]) {
  R s0 = 0, s1 = 0, s2 = 0;
  for (int v = 0; v < nv; v++) {
    s0 += X[(ci * nv + v) * 3 + 0];
    s1 += X[(ci * nv + v) * 3 + 1];
    s2 += X[(ci * nv + v) * 3 + 2];
  }
  R inv = (R)1.0f / (R)nv;
  c[0] = s0 * inv; c[1] = s1 * inv; c[2] = s2 * inv;
}

/* getVolume, Cell3D_Kernel.cl:46-64. */
static R FN(volume3)(const int* faces, int nf, const R* X, int nv, int ci) {
  R vol = 0.0f;
  for (int f = 0; f < nf; f++) {
    const R* P0 = &X[(ci * nv + faces[f * 3 + 0]) * 3];
    const R* P1 = &X[(ci * nv + faces[f * 3 + 1]) * 3];
    const R* P2 = &X[(ci * nv + faces[f * 3 + 2]) * 3];
    R c[3];
    FN(cross3)(P0, P1, c);
    vol += FN(dot3)(c, P2) / (R)6.0f;
  }
  return FABS(vol);
}

/* Solid angle sum of cell cj (shifted by `shift`) seen from point p,
 * RepellingForces Cell3D_Kernel.cl:283-300.  Returns totalOmega. */
static R FN(omega_sum)(const int* faces, int nf, const R* X, int nv, int cj,
                       const R shift[3], const R p[3]) {
  R total = 0.0f;
  for (int f = 0; f < nf; f++) {
    R a[3], b[3], c[3], u[3], v[3], w[3], vw[3];
    const R* A = &X[(cj * nv + faces[f * 3 + 0]) * 3];
    const R* B = &X[(cj * nv + faces[f * 3 + 1]) * 3];
    const R* C = &X[(cj * nv + faces[f * 3 + 2]) * 3];
    for (int k = 0; k < 3; k++) {
      a[k] = A[k] + shift[k] - p[k];
      b[k] = B[k] + shift[k] - p[k];
      c[k] = C[k] + shift[k] - p[k];
    }
    FN(normalize3)(a, u);
    FN(normalize3)(b, v);
    FN(normalize3)(c, w);
    R denom = (R)1.0f + FN(dot3)(u, v) + FN(dot3)(v, w) + FN(dot3)(w, u);
    if (denom < 1e-8f) continue;
    FN(cross3)(v, w, vw);
    R num = FN(dot3)(u, vw);
    total += (R)2.0f * ATAN2(num, denom);
  }
  return total;
}

/* Options of the 3D step that have no reference counterpart. */
typedef struct {
  unsigned mask;  /* 1 volume, 2 surface, 4 substrate, 8 contact,
                     16 vertex-vertex adhesion, 32 bending */
  R ks_grad;      /* substrate stiffness gradient along x (0 = uniform) */
} FN(opts3d);

/* One evaluation of the 3D forces, in the reference's launch order
 * Volume, SurfaceArea, StickToSurface, Repelling (Tissue3D.cpp:374-414),
 * then the two terms the reference does not launch: vertex-vertex adhesion
 * (AllVertAttraction, Cell3D_Kernel.cl:313-364) and bending.  Face
 * contributions are scattered in face order, one valid order of the
 * reference's atomics.
 * pairs (optional, nv*ncells*ncells bytes): contact pair flags
 *   pairs[(ci*nv+v)*ncells + cj]
 * wnum (optional, same shape, R): winding numbers of every evaluated pair
 * down (optional, ncells*nf bytes): faces that pass the facing-down test
 * adh_count (optional, ncells*nv ints): adhesion partners (own side)
 * adh_hash (optional, ncells*nv*2): per vertex a, the key sum (vkey3) of the
 *   partners b its own loop accepts, then of the vertices b whose own loop
 *   accepts a (the reverse pairs whose force lands on a) */
int FN(ora3d_forces)(int ncells, int nv, int nf, const int* faces,
                     const int* adj, int maxdeg, const R* X, const R* Kv,
                     const R* Ka, const R* Ks, const R* Kb, const R* v0,
                     const R* a0, R L, int PBC, R Kre, R Kat, unsigned mask,
                     R ks_grad, R* F, unsigned char* pairs, R* wnum,
                     unsigned char* down, int* adh_count,
                     unsigned* adh_hash) {
  const size_t nvt = (size_t)ncells * nv;
  for (size_t k = 0; k < nvt * 3; k++) F[k] = 0;
  if (adh_count)
    for (size_t k = 0; k < nvt; k++) adh_count[k] = 0;
  if (adh_hash)
    for (size_t k = 0; k < 2 * nvt; k++) adh_hash[k] = 0;
  R* l0 = (R*)malloc(sizeof(R) * ncells);
  R* com = (R*)malloc(sizeof(R) * ncells * 3);
  if (!l0 || !com) return -1;
  for (int ci = 0; ci < ncells; ci++) {
    /* Tissue3D.cpp:177 */
    l0[ci] = SQRT((R)4.0f * a0[ci]) / SQRT((R)3.0f);
    FN(com3)(X, nv, ci, &com[ci * 3]);
  }

  /* VolumeForceUpdate, Cell3D_Kernel.cl:66-112 */
  if (mask & 1u) {
    for (int ci = 0; ci < ncells; ci++) {
      if (Kv[ci] == 0.0f) continue;
      R volume = FN(volume3)(faces, nf, X, nv, ci);
      R volumeStrain = (volume / v0[ci]) - 1.0f;
      const R* C = &com[ci * 3];
      for (int f = 0; f < nf; f++) {
        int i0 = ci * nv + faces[f * 3 + 0];
        int i1 = ci * nv + faces[f * 3 + 1];
        int i2 = ci * nv + faces[f * 3 + 2];
        R A[3], B[3], Cc[3], g0[3], g1[3], g2[3];
        for (int k = 0; k < 3; k++) {
          A[k] = X[i1 * 3 + k] - C[k];
          B[k] = X[i2 * 3 + k] - C[k];
          Cc[k] = X[i0 * 3 + k] - C[k];
        }
        FN(cross3)(A, B, g0);
        FN(cross3)(B, Cc, g1);
        FN(cross3)(Cc, A, g2);
        R pre = -Kv[ci] * volumeStrain;
        for (int k = 0; k < 3; k++) {
          F[i0 * 3 + k] += pre * g0[k] / (R)6.0f;
          F[i1 * 3 + k] += pre * g1[k] / (R)6.0f;
          F[i2 * 3 + k] += pre * g2[k] / (R)6.0f;
        }
      }
    }
  }

  /* SurfaceAreaForceUpdate, Cell3D_Kernel.cl:114-177 */
  if (mask & 2u) {
    for (int ci = 0; ci < ncells; ci++) {
      if (Ka[ci] < 1e-8f) continue;
      for (int f = 0; f < nf; f++) {
        int idx[3] = {ci * nv + faces[f * 3 + 0], ci * nv + faces[f * 3 + 1],
                      ci * nv + faces[f * 3 + 2]};
        R lv[3][3], len[3], un[3][3], d[3];
        for (int k = 0; k < 3; k++) {
          lv[0][k] = X[idx[1] * 3 + k] - X[idx[0] * 3 + k];
          lv[1][k] = X[idx[2] * 3 + k] - X[idx[1] * 3 + k];
          lv[2][k] = X[idx[0] * 3 + k] - X[idx[2] * 3 + k];
        }
        for (int e = 0; e < 3; e++) len[e] = SQRT(FN(dot3)(lv[e], lv[e]));
        if (len[0] < 1e-12f || len[1] < 1e-12f || len[2] < 1e-12f) continue;
        for (int e = 0; e < 3; e++)
          for (int k = 0; k < 3; k++) un[e][k] = lv[e][k] / len[e];
        R target = l0[ci];
        for (int e = 0; e < 3; e++) d[e] = (len[e] / target) - 1.0f;
        R scale = Ka[ci] * SQRT(a0[ci]) / target * 0.3f;
        for (int k = 0; k < 3; k++) {
          R f0 = (un[0][k] * d[0]) - (un[2][k] * d[2]);
          R f1 = (un[1][k] * d[1]) - (un[0][k] * d[0]);
          R f2 = (un[2][k] * d[2]) - (un[1][k] * d[1]);
          F[idx[0] * 3 + k] += f0 * scale;
          F[idx[1] * 3 + k] += f1 * scale;
          F[idx[2] * 3 + k] += f2 * scale;
        }
      }
    }
  }

  /* StickToSurface, Cell3D_Kernel.cl:180-247 */
  if (down)
    for (int k = 0; k < ncells * nf; k++) down[k] = 0;
  if (mask & 4u) {
    for (int ci = 0; ci < ncells; ci++) {
      if (Ks[ci] < 1e-12f) continue;
      for (int f = 0; f < nf; f++) {
        int idx[3] = {ci * nv + faces[f * 3 + 0], ci * nv + faces[f * 3 + 1],
                      ci * nv + faces[f * 3 + 2]};
        const R* P0 = &X[idx[0] * 3];
        const R* P1 = &X[idx[1] * 3];
        const R* P2 = &X[idx[2] * 3];
        R A[3], B[3], nrm[3], un[3];
        for (int k = 0; k < 3; k++) { A[k] = P1[k] - P0[k]; B[k] = P2[k] - P0[k]; }
        FN(cross3)(A, B, nrm);
        FN(normalize3)(nrm, un);
        if (!(un[2] < -0.1f)) continue;
        if (down) down[ci * nf + f] = 1;
        const R* C = &com[ci * 3];
        for (int i = 0; i < 3; i++) {
          const R* pos = &X[idx[i] * 3];
          R ks = Ks[ci];
          if (ks_grad != (R)0) {
            R xw = pos[0] - L * FLOOR(pos[0] / L);
            ks = Ks[ci] * ((R)1 + ks_grad * xw / L);
          }
          if (pos[2] < 0.0f) F[idx[i] * 3 + 2] += ks * FABS(pos[2]);
          R height = FABS(pos[2]);
          if (height < l0[ci] * 2.0f) {
            R c2v[3] = {pos[0] - C[0], pos[1] - C[1], (R)0.0f - C[2]};
            R dz[3] = {(R)0, (R)0, pos[2] - (R)0.0f};
            R dist = SQRT(FN(dot3)(dz, dz));
            R ftmp = ks * (1.0 - dist / l0[ci]);
            R u[3];
            FN(normalize3)(c2v, u);
            for (int k = 0; k < 3; k++) F[idx[i] * 3 + k] += u[k] * ftmp;
          }
        }
      }
    }
  }

  /* RepellingForces, Cell3D_Kernel.cl:251-310 */
  if (pairs) memset(pairs, 0, nvt * (size_t)ncells);
  if (wnum)
    for (size_t k = 0; k < nvt * (size_t)ncells; k++) wnum[k] = 0;
  if ((mask & 8u) && Kre != 0.0f) {
    const R fourpi = (R)4.0f * (R)3.14159265358979323846;
    ORA_PARALLEL_FOR
    for (int ci = 0; ci < ncells; ci++) {
      const R* Ci = &com[ci * 3];
      for (int v = 0; v < nv; v++) {
        size_t gv = (size_t)ci * nv + v;
        const R* p = &X[gv * 3];
        for (int cj = 0; cj < ncells; cj++) {
          if (cj == ci) continue;
          R shift[3] = {0, 0, 0};
          if (PBC)
            for (int k = 0; k < 3; k++)
              shift[k] = L * ROUND((Ci[k] - com[cj * 3 + k]) / L);
          R tot = FN(omega_sum)(faces, nf, X, nv, cj, shift, p);
          R wn = tot / fourpi;
          if (wnum) wnum[gv * ncells + cj] = wn;
          if (FABS(wn) < 1e-6f) continue;
          if (pairs) pairs[gv * ncells + cj] = 1;
          R dv[3] = {Ci[0] - p[0], Ci[1] - p[1], Ci[2] - p[2]}, dir[3];
          FN(normalize3)(dv, dir);
          R mag = FABS(wn) * 0.5f * Kre;
          for (int k = 0; k < 3; k++) F[gv * 3 + k] += mag * dir[k];
        }
      }
    }
  }

  /* AllVertAttraction, Cell3D_Kernel.cl:313-364 (not launched by the
   * reference host; enabled here by mask bit 16). */
  if ((mask & 16u) && Kat != 0.0f) {
    for (int ci = 0; ci < ncells; ci++) {
      R L0 = l0[ci];
      if (L0 < 1e-12f) L0 = 1e-12f;
      for (int vi = 0; vi < nv; vi++) {
        size_t a = (size_t)ci * nv + vi;
        for (int cj = 0; cj < ncells; cj++) {
          if (cj == ci) continue;
          for (int vj = 0; vj < nv; vj++) {
            size_t b = (size_t)cj * nv + vj;
            R d[3];
            for (int k = 0; k < 3; k++) d[k] = X[b * 3 + k] - X[a * 3 + k];
            if (PBC)
              for (int k = 0; k < 3; k++) d[k] -= L * ROUND(d[k] / L);
            R dist = SQRT(FN(dot3)(d, d));
            if (dist < L0 * 2.0f && dist > 1e-12f) {
              for (int k = 0; k < 3; k++) {
                R t = Kat * 0.5f * (dist / L0 - 1.0f) * (d[k] / dist);
                F[a * 3 + k] += -t;
                F[b * 3 + k] += t;
              }
              if (adh_count) adh_count[a]++;
              if (adh_hash) {
                adh_hash[2 * a] += FN(vkey3)(X[b * 3], X[b * 3 + 1], X[b * 3 + 2]);
                adh_hash[2 * b + 1] += FN(vkey3)(X[a * 3], X[a * 3 + 1], X[a * 3 + 2]);
              }
            }
          }
        }
      }
    }
  }

  /* Bending (no reference counterpart): F = -Kb * L(L x) with L the
   * unnormalised graph Laplacian of the mesh edge graph, so that on a cycle
   * graph it is the reference 2D bending Kb(2s_i - s_{i-1} - s_{i+1}),
   * Cell2D_kernel.cl:80-89.  adj[v*maxdeg + k] lists the neighbours of v in
   * ascending order, -1 padded. */
  if ((mask & 32u) && adj) {
    R* s = (R*)malloc(sizeof(R) * nv * 3);
    if (!s) return -1;
    for (int ci = 0; ci < ncells; ci++) {
      if (Kb[ci] == (R)0) continue;
      for (int v = 0; v < nv; v++) {
        R acc[3] = {0, 0, 0};
        const R* xv = &X[((size_t)ci * nv + v) * 3];
        for (int k = 0; k < maxdeg; k++) {
          int u = adj[v * maxdeg + k];
          if (u < 0) break;
          const R* xu = &X[((size_t)ci * nv + u) * 3];
          for (int c = 0; c < 3; c++) acc[c] += xu[c] - xv[c];
        }
        for (int c = 0; c < 3; c++) s[v * 3 + c] = acc[c];
      }
      for (int v = 0; v < nv; v++) {
        R acc[3] = {0, 0, 0};
        int deg = 0;
        for (int k = 0; k < maxdeg; k++) {
          int u = adj[v * maxdeg + k];
          if (u < 0) break;
          deg++;
          for (int c = 0; c < 3; c++) acc[c] += s[u * 3 + c];
        }
        for (int c = 0; c < 3; c++)
          F[((size_t)ci * nv + v) * 3 + c] += -Kb[ci] * (acc[c] - (R)deg * s[v * 3 + c]);
      }
    }
    free(s);
  }
  free(l0);
  free(com);
  return 0;
}

/* Tissue3D::CLEulerUpdate's step loop, Tissue3D.cpp:372-445: clear, forces,
 * EulerPosition (Cell3D_Kernel.cl:371-381).  Fout = forces of the last step. */
int FN(ora3d_run)(int ncells, int nv, int nf, const int* faces, const int* adj,
                  int maxdeg, R* X, const R* Kv, const R* Ka, const R* Ks,
                  const R* Kb, const R* v0, const R* a0, R L, int PBC, R Kre,
                  R Kat, unsigned mask, R ks_grad, int nsteps, R dt, R* Fout) {
  size_t n3 = (size_t)ncells * nv * 3;
  R* F = (R*)calloc(n3, sizeof(R));
  if (!F) return -1;
  for (int s = 0; s < nsteps; s++) {
    int rc = FN(ora3d_forces)(ncells, nv, nf, faces, adj, maxdeg, X, Kv, Ka, Ks,
                              Kb, v0, a0, L, PBC, Kre, Kat, mask, ks_grad, F,
                              NULL, NULL, NULL, NULL, NULL);
    if (rc) { free(F); return rc; }
    for (size_t k = 0; k < n3; k++) X[k] = FMA(F[k], dt, X[k]);
  }
  if (Fout) memcpy(Fout, F, sizeof(R) * n3);
  free(F);
  return 0;
}

/* Per-cell volume (getVolume), exposed for tests. */
void FN(ora3d_volumes)(int ncells, int nv, int nf, const int* faces,
                       const R* X, R* out) {
  for (int ci = 0; ci < ncells; ci++) out[ci] = FN(volume3)(faces, nf, X, nv, ci);
}

/* Winding number of point p about cell cj (no shift), exposed for tests. */
R FN(ora3d_winding)(int nv, int nf, const int* faces, const R* X, int cj,
                    const R* p) {
  R shift[3] = {0, 0, 0};
  return FN(omega_sum)(faces, nf, X, nv, cj, shift, p) /
         ((R)4.0f * (R)3.14159265358979323846);
}

/* Elastic energy of each cell, no reference counterpart (the reference
 * computes forces only).  The energy whose negative gradient is the volume
 * force (Cell3D_Kernel.cl:66-112: F = -Kv (V/v0 - 1) dV/dx), the edge
 * springs of the surface force (:114-177: each face edge pulls with
 * sc (l/l0 - 1), sc = 0.3 Ka sqrt(a0)/l0) and the graph-Laplacian bending
 * (F = -Kb L L x):
 *   U = Kv v0/2 (V/v0 - 1)^2 + sum_faces sum_edges sc l0/2 (l/l0 - 1)^2
 *       + Kb/2 sum_v |(L x)_v|^2,
 * restricted to mask bits 1, 2, 32.  Pinned by central differences of U
 * against ora3d_forces in the tests. */
void FN(ora3d_energy)(int ncells, int nv, int nf, const int* faces,
                      const int* adj, int maxdeg, const R* X, const R* Kv,
                      const R* Ka, const R* Kb, const R* v0, const R* a0,
                      unsigned mask, R* E) {
  for (int ci = 0; ci < ncells; ci++) {
    const R l0 = SQRT((R)4.0f * a0[ci]) / SQRT((R)3.0f);
    R e = 0;
    if ((mask & 1u) && Kv[ci] != 0.0f) {
      const R s = FN(volume3)(faces, nf, X, nv, ci) / v0[ci] - (R)1;
      e += (R)0.5 * Kv[ci] * v0[ci] * s * s;
    }
    if ((mask & 2u) && !(Ka[ci] < 1e-8f)) {
      const R sc = Ka[ci] * SQRT(a0[ci]) / l0 * 0.3f;
      for (int f = 0; f < nf; f++)
        for (int k = 0; k < 3; k++) {
          const R* P = &X[((size_t)ci * nv + faces[f * 3 + k]) * 3];
          const R* Q = &X[((size_t)ci * nv + faces[f * 3 + (k + 1) % 3]) * 3];
          const R d[3] = {Q[0] - P[0], Q[1] - P[1], Q[2] - P[2]};
          const R s = SQRT(FN(dot3)(d, d)) / l0 - (R)1;
          e += (R)0.5 * sc * l0 * s * s;
        }
    }
    if ((mask & 32u) && Kb[ci] != 0.0f) {
      for (int v = 0; v < nv; v++) {
        R s[3] = {0, 0, 0};
        for (int k = 0; k < maxdeg; k++) {
          const int u = adj[v * maxdeg + k];
          if (u < 0) break;
          for (int c = 0; c < 3; c++)
            s[c] += X[((size_t)ci * nv + u) * 3 + c] - X[((size_t)ci * nv + v) * 3 + c];
        }
        e += (R)0.5 * Kb[ci] * FN(dot3)(s, s);
      }
    }
    E[ci] = e;
  }
}
