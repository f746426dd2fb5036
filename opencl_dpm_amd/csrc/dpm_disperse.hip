// Dispersal of cell centres on the device (Tissue2D::Disperse,
// src/Tissue2D.cpp:42-105, and Tissue3D::Disperse2D, src/Tissue3D.cpp:40-106).
//
// The reference relaxes soft disks in a periodic box on the host, O(n^2)
// per iteration and up to 100001 iterations: every ordered pair (i, j)
// closer than r_i + r_j pushes both centres apart; the iteration stops when
// the float energy sum U changes by at most 1e-6.  Its outcome depends on
// every rounding: the float force sums are accumulated in the loop order of
// (i, j), and U in that order too.  This kernel reproduces them bit for bit:
//   * one workgroup holds all centres in LDS and runs every iteration;
//   * centre k's force is the sum, in the reference's order, of the terms
//     that reach it: +f(i, k) for contact partners i < k, then -f(k, j) for
//     its own partners j ascending, then +f(i, k) for partners i > k (pairs
//     out of contact add nothing, so only partners are visited);
//   * partners come from a periodic bin grid of width >= the largest contact
//     distance (3 x 3 bins around the centre), gathered and sorted by index;
//     a centre with more than kDispMaxCand candidates, or a box with fewer
//     than 3 bins per axis, scans every centre in order instead;
//   * U's terms are written in (i, j) order and one lane adds them in
//     float as the reference does.
// The float/double mix of each expression is the reference's (double
// literals promote), and every source is compiled with -ffp-contract=off.
#pragma once

constexpr int kDispMaxCand = 48;       // gathered candidates per centre
constexpr int kDispThreads = 1024;  // at most; the launch uses about one thread per centre

struct Disperse2DArgs {
  const float* r;  // per centre: radius * scale
  float* X;        // in: starting centres; out: relaxed
  float* Y;
  int n;
  float L;
  int max_count;   // the loop stops once it has run more than this many iterations
  int nb;          // bins per axis (0: every centre scans all)
  int* out;        // [0] iterations run, [1] 1 if stopped by max_count
  float* terms;    // [term_cap] U terms in (i, j) order
  int term_cap;
};

struct DispLds {
  float *x, *y, *r, *fx, *fy;
  int* cnt;        // [n + 1] own-row contact counts, then their exclusive scan
  int* bstart;     // [nb*nb + 1]
  int* blist;      // [n]
  int* bfill;      // [nb*nb]
  unsigned short* cand;  // [threads * kDispMaxCand]
  float* red;      // [4] U, oldU, dU, flags
  size_t total;
};

__host__ __device__ inline DispLds disp_layout(unsigned char* base, int n, int nb, int threads) {
  DispLds l;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    unsigned char* p = base ? base + o : nullptr;
    o = (o + bytes + 15) & ~(size_t)15;
    return p;
  };
  l.x = (float*)take(4u * n);
  l.y = (float*)take(4u * n);
  l.r = (float*)take(4u * n);
  l.fx = (float*)take(4u * n);
  l.fy = (float*)take(4u * n);
  l.cnt = (int*)take(4u * (n + 1));
  const int nbin = nb > 0 ? nb * nb : 0;
  l.bstart = (int*)take(4u * (nbin + 1));
  l.blist = (int*)take(4u * n);
  l.bfill = (int*)take(4u * (nbin + 1));
  l.cand = (unsigned short*)take(2u * (size_t)threads * kDispMaxCand);
  l.red = (float*)take(16u);
  l.total = o;
  return l;
}

// The pair term of centre a on centre b's coordinates (the reference's row a,
// column b): whether they touch, and the force component pushed onto b
// (row a adds it to b and subtracts it from a), and the U term.
struct PairTerm {
  bool touch;
  float fx, fy, u;
};

__device__ __forceinline__ PairTerm disp_pair(float xa, float ya, float ra, float xb, float yb, float rb, float L) {
  PairTerm t{false, 0.0f, 0.0f, 0.0f};
  float dx = xb - xa;
  dx -= L * roundf(dx / L);
  float dy = yb - ya;
  dy -= L * roundf(dy / L);
  const float dist = sqrtf(dx * dx + dy * dy);
  const float rr = ra + rb;
  if (dist <= rr) {
    const float ux = dx / dist, uy = dy / dist;
    const float q = dist / rr;
    const float ftmp = (float)((1.0 - (double)q) / (double)rr);
    t.touch = true;
    t.fx = ftmp * ux;
    t.fy = ftmp * uy;
    t.u = 1.0f - q * (1.0f - q);
  }
  return t;
}

__device__ __forceinline__ int disp_bin(float x, float L, float w, int nb) {
  const float xw = x - L * floorf(x / L);
  const int b = (int)floorf(xw / w);
  return b < 0 ? 0 : (b >= nb ? nb - 1 : b);
}

// Candidates of centre k, ascending, into c[]; returns their number, or -1
// when the scan of every centre applies.
__device__ __forceinline__ int disp_gather(const DispLds& l, int k, int n, int nb, float L, unsigned short* c) {
  if (nb < 3) return -1;
  const float w = L / (float)nb;
  const int bx = disp_bin(l.x[k], L, w, nb), by = disp_bin(l.y[k], L, w, nb);
  int m = 0;
  for (int dy = -1; dy <= 1; dy++)
    for (int dx = -1; dx <= 1; dx++) {
      const int b = ((by + dy + nb) % nb) * nb + (bx + dx + nb) % nb;
      for (int q = l.bstart[b]; q < l.bstart[b + 1]; q++) {
        const int i = l.blist[q];
        if (i == k) continue;
        if (m == kDispMaxCand) return -1;
        c[m++] = (unsigned short)i;
      }
    }
  for (int a = 1; a < m; a++) {  // insertion sort (indices are distinct)
    const unsigned short v = c[a];
    int b = a - 1;
    while (b >= 0 && c[b] > v) {
      c[b + 1] = c[b];
      b--;
    }
    c[b + 1] = v;
  }
  return m;
}

extern "C" __global__ void __launch_bounds__(kDispThreads) k_disperse2d(const Disperse2DArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = a.n, nb = a.nb, tid = threadIdx.x, NT = blockDim.x;
  const float L = a.L;
  const DispLds l = disp_layout(smem, n, nb, NT);
  unsigned short* mycand = l.cand + (size_t)tid * kDispMaxCand;
  for (int k = tid; k < n; k += NT) {
    l.x[k] = a.X[k];
    l.y[k] = a.Y[k];
    l.r[k] = a.r[k];
  }
  if (tid == 0) {
    l.red[1] = 100.0f;  // oldU
    l.red[2] = 100.0f;  // dU
  }
  __syncthreads();
  int count = 0;
  bool capped = false;
  // while (dU > 1e-6): the loop condition in double, as the reference's
  // float dU against a double literal
  while ((double)l.red[2] > 1e-6) {
    // ---- bins of the current centres (counting sort)
    if (nb >= 3) {
      const int nbin = nb * nb;
      const float w = L / (float)nb;
      for (int b = tid; b <= nbin; b += NT) l.bfill[b] = 0;
      __syncthreads();
      for (int k = tid; k < n; k += NT) atomicAdd(&l.bfill[disp_bin(l.y[k], L, w, nb) * nb + disp_bin(l.x[k], L, w, nb)], 1);
      __syncthreads();
      if (tid == 0) {
        int run = 0;
        for (int b = 0; b < nbin; b++) {
          l.bstart[b] = run;
          run += l.bfill[b];
          l.bfill[b] = l.bstart[b];
        }
        l.bstart[nbin] = run;
      }
      __syncthreads();
      for (int k = tid; k < n; k += NT) {
        const int b = disp_bin(l.y[k], L, w, nb) * nb + disp_bin(l.x[k], L, w, nb);
        l.blist[atomicAdd(&l.bfill[b], 1)] = k;  // order within a bin: sorted per centre below
      }
      __syncthreads();
    }
    // ---- forces on every centre, in the reference's accumulation order
    int m_own = -2;  // this thread's gathered list, kept for the U terms when n <= NT
    for (int k = tid; k < n; k += NT) {
      const int m = disp_gather(l, k, n, nb, L, mycand);
      m_own = m;
      const int ncand = m < 0 ? n : m;
      const float xk = l.x[k], yk = l.y[k], rk = l.r[k];
      float Fx = 0.0f, Fy = 0.0f;
      int own = 0;
      // rows i < k, then the own row, then rows i > k
      for (int q = 0; q < ncand; q++) {
        const int i = m < 0 ? q : (int)mycand[q];
        if (i >= k) break;
        const PairTerm t = disp_pair(l.x[i], l.y[i], l.r[i], xk, yk, rk, L);
        if (t.touch) {
          Fx += t.fx;
          Fy += t.fy;
        }
      }
      for (int q = 0; q < ncand; q++) {
        const int j = m < 0 ? q : (int)mycand[q];
        if (j == k) continue;
        const PairTerm t = disp_pair(xk, yk, rk, l.x[j], l.y[j], l.r[j], L);
        if (t.touch) {
          Fx -= t.fx;
          Fy -= t.fy;
          own++;
        }
      }
      for (int q = 0; q < ncand; q++) {
        const int i = m < 0 ? q : (int)mycand[q];
        if (i <= k) continue;
        const PairTerm t = disp_pair(l.x[i], l.y[i], l.r[i], xk, yk, rk, L);
        if (t.touch) {
          Fx += t.fx;
          Fy += t.fy;
        }
      }
      l.fx[k] = Fx;
      l.fy[k] = Fy;
      l.cnt[k] = own;
    }
    __syncthreads();
    if (tid == 0) {  // exclusive scan of the rows' term counts
      int run = 0;
      for (int k = 0; k < n; k++) {
        const int c = l.cnt[k];
        l.cnt[k] = run;
        run += c;
      }
      l.cnt[n] = run;
    }
    __syncthreads();
    const bool stored = l.cnt[n] <= a.term_cap;  // block-uniform
    if (stored) {  // the own rows' U terms at their places in (i, j) order
      for (int k = tid; k < n; k += NT) {
        // a thread with one centre still holds its list (positions unchanged)
        const int m = n <= NT ? m_own : disp_gather(l, k, n, nb, L, mycand);
        const int ncand = m < 0 ? n : m;
        int o = l.cnt[k];
        for (int q = 0; q < ncand; q++) {
          const int j = m < 0 ? q : (int)mycand[q];
          if (j == k) continue;
          const PairTerm t = disp_pair(l.x[k], l.y[k], l.r[k], l.x[j], l.y[j], l.r[j], L);
          if (t.touch) a.terms[o++] = t.u;
        }
      }
    }
    __syncthreads();
    if (tid == 0) {
      // U += 0.5 * (1 - q (1 - q)), in the reference's (i, j) order
      float U = 0.0f;
      if (stored) {
        // the terms in order, eight loads in flight ahead of their additions
        const int tot = l.cnt[n];
        int q = 0;
        for (; q + 8 <= tot; q += 8) {
          float tb[8];
#pragma unroll
          for (int u = 0; u < 8; u++) tb[u] = a.terms[q + u];
#pragma unroll
          for (int u = 0; u < 8; u++) U = (float)((double)U + 0.5 * (double)tb[u]);
        }
        for (; q < tot; q++) U = (float)((double)U + 0.5 * (double)a.terms[q]);
      } else {
        for (int i = 0; i < n; i++)
          for (int j = 0; j < n; j++) {
            if (j == i) continue;
            const PairTerm t = disp_pair(l.x[i], l.y[i], l.r[i], l.x[j], l.y[j], l.r[j], L);
            if (t.touch) U = (float)((double)U + 0.5 * (double)t.u);
          }
      }
      l.red[0] = U;
    }
    __syncthreads();  // every read of the old centres is done
    for (int k = tid; k < n; k += NT) {  // X[i] += 0.01 * Fx[i], in double
      l.x[k] = (float)((double)l.x[k] + 0.01 * (double)l.fx[k]);
      l.y[k] = (float)((double)l.y[k] + 0.01 * (double)l.fy[k]);
    }
    if (tid == 0) {
      float dU = l.red[0] - l.red[1];
      if ((double)dU < 0.0) dU *= -1.0f;
      l.red[1] = l.red[0];
      l.red[2] = dU;
    }
    count++;
    if (count > a.max_count) {  // block-uniform
      capped = true;
      __syncthreads();
      break;
    }
    __syncthreads();
  }
  for (int k = tid; k < n; k += NT) {
    a.X[k] = l.x[k];
    a.Y[k] = l.y[k];
  }
  if (tid == 0) {
    a.out[0] = count;
    a.out[1] = capped ? 1 : 0;
  }
}
