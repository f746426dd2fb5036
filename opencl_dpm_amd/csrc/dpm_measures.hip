// The reference's per-cell host reductions on the device.
//
// Tissue3D::CLEulerUpdate ends each call by recomputing every cell's volume
// and surface area on the host (src/Tissue3D.cpp:503-504) with
// Cell3D::GetVolume and Cell3D::GetSurfaceArea (src/cell.cpp:196-230), and
// Disperse2D centres cells with Cell3D::GetCOM (src/cell.cpp:276-288).  All
// three are serial float sums over faces or vertices.  Here one workgroup per
// cell stages the cell (positions rounded to float, faces) in LDS and one
// lane runs the three sums in the host's operation order, so the results
// are bit-identical to the host functions (every source is compiled with
// -ffp-contract=off and correctly rounded float division and square root,
// the host code has no fused multiply-add).
#pragma once

template <typename R>
struct Measures3D {
  const R* x;
  const R* y;
  const R* z;
  const int* faces;  // [nf*3]
  int nv, nf;
  float* out;  // per cell: volume, surface area, centre of mass x, y, z
};

// LDS: 3 nv floats (x | y | z), then 3 nf ints.
inline size_t measures_lds_bytes(int nv, int nf) { return (size_t)3 * nv * 4 + (size_t)3 * nf * 4; }

template <typename R>
__device__ void k3d_measures_body(const Measures3D<R>& a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int c = blockIdx.x, nv = a.nv, nf = a.nf;
  float* px = reinterpret_cast<float*>(smem);
  float* py = px + nv;
  float* pz = py + nv;
  int* fc = reinterpret_cast<int*>(pz + nv);
  const size_t base = (size_t)c * nv;
  for (int v = threadIdx.x; v < nv; v += blockDim.x) {
    px[v] = (float)a.x[base + v];
    py[v] = (float)a.y[base + v];
    pz[v] = (float)a.z[base + v];
  }
  for (int k = threadIdx.x; k < 3 * nf; k += blockDim.x) fc[k] = a.faces[k];
  __syncthreads();
  if (threadIdx.x != 0) return;
  // GetVolume (cell.cpp:196-211): the triple product p0 . (p1 x p2) per face,
  // summed in float, |sum| / 6 divided in double and returned as float.
  float volume = 0.0f, sa = 0.0f;
  for (int f = 0; f < nf; f++) {
    const int i0 = fc[3 * f], i1 = fc[3 * f + 1], i2 = fc[3 * f + 2];
    const float p0x = px[i0], p0y = py[i0], p0z = pz[i0];
    const float p1x = px[i1], p1y = py[i1], p1z = pz[i1];
    const float p2x = px[i2], p2y = py[i2], p2z = pz[i2];
    const float c0 = p1y * p2z - p1z * p2y;
    const float c1 = p1z * p2x - p1x * p2z;
    const float c2 = p1x * p2y - p1y * p2x;
    volume += (c0 * p0x + c1 * p0y) + c2 * p0z;
    // GetSurfaceArea (cell.cpp:213-230): |(p1 - p0) x (p2 - p0)| per face
    const float A0 = p1x - p0x, A1 = p1y - p0y, A2 = p1z - p0z;
    const float B0 = p2x - p0x, B1 = p2y - p0y, B2 = p2z - p0z;
    const float d0 = A1 * B2 - A2 * B1;
    const float d1 = A2 * B0 - A0 * B2;
    const float d2 = A0 * B1 - A1 * B0;
    sa += sqrtf((d0 * d0 + d1 * d1) + d2 * d2);
  }
  // GetCOM (cell.cpp:276-288): per-component float sums, then / (float)NV
  float cx = 0.0f, cy = 0.0f, cz = 0.0f;
  for (int v = 0; v < nv; v++) {
    cx += px[v];
    cy += py[v];
    cz += pz[v];
  }
  float* o = a.out + 5 * (size_t)c;
  o[0] = (float)((double)fabsf(volume) / 6.0);
  o[1] = sa;
  o[2] = cx / (float)nv;
  o[3] = cy / (float)nv;
  o[4] = cz / (float)nv;
}

extern "C" __global__ void __launch_bounds__(64) k3d_measures_f32(const Measures3D<float> a) { k3d_measures_body(a); }
extern "C" __global__ void __launch_bounds__(64) k3d_measures_f64(const Measures3D<double> a) { k3d_measures_body(a); }
