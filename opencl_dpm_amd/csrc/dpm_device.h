// Device helpers shared by the 2D and 3D DPM kernels (gfx950, wave64).
//
// Arithmetic contract: every kernel source is compiled with
// -ffp-contract=off and HIP's correctly rounded float divide and sqrt, so
// each + - * / sqrt below is one IEEE-754 operation.  The expressions that
// decide a pair or a face (winding denominators, crossing tests, distance
// cut-offs, the facing-down test) are written in the same operation order
// as the reference kernels, which makes those decisions bit-identical to a
// CPU evaluation of the same order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define DPM_WAVE 64
// Diagnostic counters are kept in this many slots of 4 (summed on the host).
#define DPM_COUNTER_SLOTS 256

namespace dpm {

struct f3 {
  float x, y, z;
};

__device__ __forceinline__ f3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 add(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 scale(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
// dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z, unfused.
__device__ __forceinline__ float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// normalize(v) = v * (1/sqrt(dot(v,v))), normalize(0) = 0.
__device__ __forceinline__ f3 normalize(f3 a) {
  float d = dot(a, a);
  if (d == 0.0f) return {0.0f, 0.0f, 0.0f};
  float inv = 1.0f / sqrtf(d);
  return {a.x * inv, a.y * inv, a.z * inv};
}
__device__ __forceinline__ void normalize2(float x, float y, float& ux, float& uy) {
  float d = x * x + y * y;
  if (d == 0.0f) { ux = 0.0f; uy = 0.0f; return; }
  float inv = 1.0f / sqrtf(d);
  ux = x * inv;
  uy = y * inv;
}

// ------------------------------------------- the same helpers for float/double
// (the 3D step is instantiated in both; double evaluates the same formulas
// with the float thresholds promoted, as the FP64 restatement does)
template <typename R>
struct V3 {
  R x, y, z;
};
template <typename R>
__device__ __forceinline__ V3<R> v3(R x, R y, R z) { return {x, y, z}; }
template <typename R>
__device__ __forceinline__ V3<R> vsub(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R>
__device__ __forceinline__ V3<R> vadd(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R>
__device__ __forceinline__ V3<R> vscale(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename R>
__device__ __forceinline__ R vdot(V3<R> a, V3<R> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <typename R>
__device__ __forceinline__ V3<R> vcross(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// Fused forms for force arithmetic that decides no cut-off (the sources are
// compiled without contraction so that every decision keeps the reference's
// rounding; these fuse where it does not matter).
template <typename R>
__device__ __forceinline__ R vdot_f(V3<R> a, V3<R> b) { return fma(a.x, b.x, fma(a.y, b.y, a.z * b.z)); }
// c + a s
template <typename R>
__device__ __forceinline__ V3<R> vmadd_f(V3<R> a, R s, V3<R> c) {
  return {fma(a.x, s, c.x), fma(a.y, s, c.y), fma(a.z, s, c.z)};
}
// c + a x b
template <typename R>
__device__ __forceinline__ V3<R> vcross_add_f(V3<R> a, V3<R> b, V3<R> c) {
  return {fma(a.y, b.z, fma(-a.z, b.y, c.x)), fma(a.z, b.x, fma(-a.x, b.z, c.y)), fma(a.x, b.y, fma(-a.y, b.x, c.z))};
}
template <typename R>
__device__ __forceinline__ V3<R> vnormalize(V3<R> a) {
  const R d = vdot(a, a);
  if (d == R(0)) return {R(0), R(0), R(0)};
  const R inv = R(1) / sqrt(d);
  return {a.x * inv, a.y * inv, a.z * inv};
}

// ------------------------------------------------------ pair-set fingerprints
// Diagnostics of the probe entry points: a partner vertex is identified by
// the bit patterns of its coordinates, key = f(x ^ f(y ^ f(z))) with f the
// 32-bit finaliser of MurmurHash3 (a bijection of 32-bit words), and a
// vertex's partner set by the wrapping sum of its partners' keys, which does
// not depend on the order in which pairs are found (atomics may add them).
__device__ __forceinline__ unsigned fmix32(unsigned h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
__device__ __forceinline__ unsigned rbits(float v) { return __float_as_uint(v); }
__device__ __forceinline__ unsigned rbits(double v) {  // both halves of a double
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  return fmix32((unsigned)b) ^ (unsigned)(b >> 32);
}
template <typename R>
__device__ __forceinline__ unsigned vertex_key(R x, R y) {
  return fmix32(rbits(x) ^ fmix32(rbits(y)));
}
template <typename R>
__device__ __forceinline__ unsigned vertex_key(R x, R y, R z) {
  return fmix32(rbits(x) ^ fmix32(rbits(y) ^ fmix32(rbits(z))));
}

// Key of a partner cell (its index within the tissue), for the same sums.
__device__ __forceinline__ unsigned cell_key(int c) { return fmix32((unsigned)c + 0x9E3779B9u); }

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, DPM_WAVE);
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_down(v, o, DPM_WAVE));
  return v;
}

// Wave-wide maximum and minimum of a float, every lane receiving the result
// (a wave-uniform value), by DPP row operations: quad swaps, the half-row
// and row mirrors, then the row broadcasts, which leave the full result in
// lane 63 — no LDS round trip per step, as a shuffle (ds_bpermute) takes.
// Call with every lane of the wave active.  Maxima and minima are exact, so
// the result is the shuffle reduction's bit for bit.  Doubles take the
// shuffles.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, ROW_MASK, 0xF, false));
}
__device__ __forceinline__ float wave_max_f(float v) {
  v = fmaxf(v, dpp_f<0xB1, 0xF>(v));   // quad_perm(1, 0, 3, 2)
  v = fmaxf(v, dpp_f<0x4E, 0xF>(v));   // quad_perm(2, 3, 0, 1)
  v = fmaxf(v, dpp_f<0x141, 0xF>(v));  // row_half_mirror
  v = fmaxf(v, dpp_f<0x140, 0xF>(v));  // row_mirror
  v = fmaxf(v, dpp_f<0x142, 0xA>(v));  // row_bcast15 into rows 1 and 3
  v = fmaxf(v, dpp_f<0x143, 0xC>(v));  // row_bcast31 into rows 2 and 3
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ float wave_min_f(float v) {
  v = fminf(v, dpp_f<0xB1, 0xF>(v));
  v = fminf(v, dpp_f<0x4E, 0xF>(v));
  v = fminf(v, dpp_f<0x141, 0xF>(v));
  v = fminf(v, dpp_f<0x140, 0xF>(v));
  v = fminf(v, dpp_f<0x142, 0xA>(v));
  v = fminf(v, dpp_f<0x143, 0xC>(v));
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ double wave_max_all(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, DPM_WAVE));
  return v;
}
__device__ __forceinline__ float wave_max_all(float v) { return wave_max_f(v); }
__device__ __forceinline__ double wave_min_all(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, DPM_WAVE));
  return v;
}
__device__ __forceinline__ float wave_min_all(float v) { return wave_min_f(v); }

// Wave-wide sum of a double by the same DPP steps (each moves the two halves;
// a row that a row broadcast does not write adds 0), the result in every
// lane: lane 63's, read back as a wave-uniform value.  The additions are
// grouped by the DPP tree (quads, eights, rows, then rows 0+1, 2+3, all):
// a fixed order, so the sum is deterministic.  Every lane active.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_d(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, ROW_MASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, ROW_MASK, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double wave_sum_d(double v) {
  v += dpp_d<0xB1, 0xF>(v);
  v += dpp_d<0x4E, 0xF>(v);
  v += dpp_d<0x141, 0xF>(v);
  v += dpp_d<0x140, 0xF>(v);
  v += dpp_d<0x142, 0xA>(v);
  v += dpp_d<0x143, 0xC>(v);
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), 63);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Block-wide sum of one double per thread; every thread gets the result.
// scratch: >= 17 doubles of LDS.  Contains __syncthreads: call from all threads.
__device__ __forceinline__ double block_sum(double v, double* scratch) {
  const int lane = threadIdx.x & (DPM_WAVE - 1), wave = threadIdx.x / DPM_WAVE;
  const int nw = (blockDim.x + DPM_WAVE - 1) / DPM_WAVE;
  v = wave_sum(v);
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < nw; w++) s += scratch[w];
    scratch[16] = s;
  }
  __syncthreads();
  double r = scratch[16];
  __syncthreads();
  return r;
}
// Three block-wide sums in one pass (every thread gets them); scratch: >= 52
// doubles of LDS.  Contains __syncthreads.
__device__ __forceinline__ void block_sum3(double& a, double& b, double& c, double* scratch) {
  const int lane = threadIdx.x & (DPM_WAVE - 1), wave = threadIdx.x / DPM_WAVE;
  const int nw = (blockDim.x + DPM_WAVE - 1) / DPM_WAVE;
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  if (lane == 0) {
    scratch[wave] = a;
    scratch[16 + wave] = b;
    scratch[32 + wave] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sa = 0.0, sb = 0.0, sc = 0.0;
    for (int w = 0; w < nw; w++) {
      sa += scratch[w];
      sb += scratch[16 + w];
      sc += scratch[32 + w];
    }
    scratch[48] = sa;
    scratch[49] = sb;
    scratch[50] = sc;
  }
  __syncthreads();
  a = scratch[48];
  b = scratch[49];
  c = scratch[50];
  __syncthreads();
}

// block_sum3 with one workgroup barrier: every thread adds the waves'
// partial sums itself, in wave order (the same result as one thread adding
// them and broadcasting).  scratch (>= 48 doubles) must not be written by
// another reduction until every thread has passed a later barrier: callers
// alternate scratch regions.
__device__ __forceinline__ void block_sum3_1b(double& a, double& b, double& c, double* scratch) {
  const int lane = threadIdx.x & (DPM_WAVE - 1), wave = threadIdx.x / DPM_WAVE;
  const int nw = (blockDim.x + DPM_WAVE - 1) / DPM_WAVE;
  a = wave_sum_d(a);  // DPP trees (no LDS round trip per step)
  b = wave_sum_d(b);
  c = wave_sum_d(c);
  if (lane == 0) {
    scratch[wave] = a;
    scratch[16 + wave] = b;
    scratch[32 + wave] = c;
  }
  __syncthreads();
  double sa = 0.0, sb = 0.0, sc = 0.0;
  for (int w = 0; w < nw; w++) {
    sa += scratch[w];
    sb += scratch[16 + w];
    sc += scratch[32 + w];
  }
  a = sa;
  b = sb;
  c = sc;
}

// block_max_sum_max with one workgroup barrier (as block_sum3_1b).
// FLOAT_MAX: the maxima hold float values (exact as floats: DPP maxima);
// otherwise double maxima by shuffles.  The sum by the DPP tree.
template <bool FLOAT_MAX>
__device__ __forceinline__ void block_max_sum_max_1b(double& m1, double& s, double& m2, double* scratch) {
  const int lane = threadIdx.x & (DPM_WAVE - 1), wave = threadIdx.x / DPM_WAVE;
  const int nw = (blockDim.x + DPM_WAVE - 1) / DPM_WAVE;
  if constexpr (FLOAT_MAX) {
    m1 = (double)wave_max_f((float)m1);
    m2 = (double)wave_max_f((float)m2);
  } else {
    m1 = wave_max_all(m1);
    m2 = wave_max_all(m2);
  }
  s = wave_sum_d(s);
  if (lane == 0) {
    scratch[wave] = m1;
    scratch[16 + wave] = s;
    scratch[32 + wave] = m2;
  }
  __syncthreads();
  double a = scratch[0], b = 0.0, c = scratch[32];
  for (int w = 0; w < nw; w++) {
    a = fmax(a, scratch[w]);
    b += scratch[16 + w];
    c = fmax(c, scratch[32 + w]);
  }
  m1 = a;
  s = b;
  m2 = c;
}

// max, sum, max of three per-thread values in one pass (every thread gets
// them); scratch: >= 52 doubles of LDS.  Contains __syncthreads.
__device__ __forceinline__ void block_max_sum_max(double& m1, double& s, double& m2, double* scratch) {
  const int lane = threadIdx.x & (DPM_WAVE - 1), wave = threadIdx.x / DPM_WAVE;
  const int nw = (blockDim.x + DPM_WAVE - 1) / DPM_WAVE;
  for (int o = 32; o > 0; o >>= 1) {
    m1 = fmax(m1, __shfl_down(m1, o, DPM_WAVE));
    s += __shfl_down(s, o, DPM_WAVE);
    m2 = fmax(m2, __shfl_down(m2, o, DPM_WAVE));
  }
  if (lane == 0) {
    scratch[wave] = m1;
    scratch[16 + wave] = s;
    scratch[32 + wave] = m2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = scratch[0], b = 0.0, c = scratch[32];
    for (int w = 0; w < nw; w++) {
      a = fmax(a, scratch[w]);
      b += scratch[16 + w];
      c = fmax(c, scratch[32 + w]);
    }
    scratch[48] = a;
    scratch[49] = b;
    scratch[50] = c;
  }
  __syncthreads();
  m1 = scratch[48];
  s = scratch[49];
  m2 = scratch[50];
  __syncthreads();
}

// Two block-wide maxima in one pass (every thread gets them); scratch: >= 34
// values of LDS.  Contains __syncthreads: call from all threads.
template <typename R>
__device__ __forceinline__ void block_max2(R& a, R& b, R* scratch) {
  const int lane = threadIdx.x & (DPM_WAVE - 1), wave = threadIdx.x / DPM_WAVE;
  const int nw = (blockDim.x + DPM_WAVE - 1) / DPM_WAVE;
  for (int o = 32; o > 0; o >>= 1) {
    a = fmax(a, __shfl_down(a, o, DPM_WAVE));
    b = fmax(b, __shfl_down(b, o, DPM_WAVE));
  }
  if (lane == 0) {
    scratch[wave] = a;
    scratch[16 + wave] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    R x = scratch[0], y = scratch[16];
    for (int w = 1; w < nw; w++) {
      x = fmax(x, scratch[w]);
      y = fmax(y, scratch[16 + w]);
    }
    scratch[32] = x;
    scratch[33] = y;
  }
  __syncthreads();
  a = scratch[32];
  b = scratch[33];
  __syncthreads();
}

template <typename R>
__device__ __forceinline__ R block_max(R v, R* scratch) {
  const int lane = threadIdx.x & (DPM_WAVE - 1), wave = threadIdx.x / DPM_WAVE;
  const int nw = (blockDim.x + DPM_WAVE - 1) / DPM_WAVE;
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, DPM_WAVE));
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    R s = scratch[0];
    for (int w = 1; w < nw; w++) s = fmax(s, scratch[w]);
    scratch[16] = s;
  }
  __syncthreads();
  R r = scratch[16];
  __syncthreads();
  return r;
}

// Order-preserving block compaction: threads with flag set write `value`
// to list[] in thread order.  Returns the count to every thread.
// counts: >= 17 ints of LDS.  Contains __syncthreads.
__device__ __forceinline__ int block_compact(bool flag, int value, int* list, int* counts) {
  const int lane = threadIdx.x & (DPM_WAVE - 1), wave = threadIdx.x / DPM_WAVE;
  const int nw = (blockDim.x + DPM_WAVE - 1) / DPM_WAVE;
  const unsigned long long m = __ballot(flag);
  if (lane == 0) counts[wave] = __popcll(m);
  __syncthreads();
  // every thread adds the waves' counts itself (no serial pass, no third barrier)
  int off = 0, total = 0;
  for (int w = 0; w < nw; w++) {
    const int c = counts[w];
    off += w < wave ? c : 0;
    total += c;
  }
  if (flag) list[off + __popcll(m & ((1ull << lane) - 1ull))] = value;
  __syncthreads();  // the list is complete; every read of counts is done
  return total;
}

// block_compact that also calls also(slot) on every flagged thread, with
// the thread's index in the list, before the closing barrier (so what it
// writes to LDS is visible to every thread on return).
template <typename Fn>
__device__ __forceinline__ int block_compact_with(bool flag, int value, int* list, int* counts, Fn&& also) {
  const int lane = threadIdx.x & (DPM_WAVE - 1), wave = threadIdx.x / DPM_WAVE;
  const int nw = (blockDim.x + DPM_WAVE - 1) / DPM_WAVE;
  const unsigned long long m = __ballot(flag);
  if (lane == 0) counts[wave] = __popcll(m);
  __syncthreads();
  int off = 0, total = 0;
  for (int w = 0; w < nw; w++) {
    const int c = counts[w];
    off += w < wave ? c : 0;
    total += c;
  }
  if (flag) {
    const int slot = off + __popcll(m & ((1ull << lane) - 1ull));
    list[slot] = value;
    also(slot);
  }
  __syncthreads();
  return total;
}

// Two stream compactions in one pass: threads with fa append va to la,
// threads with fb append vb to lb, in thread order; slot_b is the calling
// thread's index in lb (meaningful where fb).  counts holds >= 34 ints.
__device__ __forceinline__ void block_compact2(bool fa, int va, int* la, bool fb, int vb, int* lb, int* counts,
                                               int& na, int& nb, int& slot_b) {
  const int lane = threadIdx.x & (DPM_WAVE - 1), wave = threadIdx.x / DPM_WAVE;
  const int nw = (blockDim.x + DPM_WAVE - 1) / DPM_WAVE;
  const unsigned long long ma = __ballot(fa), mb = __ballot(fb);
  const unsigned long long below = (1ull << lane) - 1ull;
  if (lane == 0) {
    counts[wave] = __popcll(ma);
    counts[16 + wave] = __popcll(mb);
  }
  __syncthreads();
  int oa = 0, ob = 0, sa = 0, sb = 0;
  for (int w = 0; w < nw; w++) {
    const int ca = counts[w], cb = counts[16 + w];
    oa += w < wave ? ca : 0;
    ob += w < wave ? cb : 0;
    sa += ca;
    sb += cb;
  }
  const int pa = oa + __popcll(ma & below);
  slot_b = ob + __popcll(mb & below);
  if (fa) la[pa] = va;
  if (fb) lb[slot_b] = vb;
  na = sa;
  nb = sb;
  __syncthreads();
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X
// (workgroup b runs on XCD b % 8).  This bijection hands each XCD a
// contiguous range of work items, so spatially neighbouring cells — which
// read each other's vertices — share one XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
  const int q = n / 8, r = n % 8;
  const int x = b % 8, i = b / 8;
  return x * q + (x < r ? x : r) + i;
}

}  // namespace dpm
