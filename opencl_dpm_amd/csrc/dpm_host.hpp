// Host-side model behind the Python module `clDPM`: the Cell2D, Cell3D,
// Tissue2D and Tissue3D classes with the reference's fields, constructors
// and methods (include/cell.hpp:11-58, include/Tissue.hpp:13-55).  The
// timestep itself is delegated to the C ABI of libdpm_hip (include/dpm.h);
// nothing here computes forces.
#pragma once
#include <array>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

struct dpm_tissue;

namespace dpmhost {

// A process-wide unique number drawn at every change of a cell's vertex
// positions (construction, assignment from Python, dispersal, a timestep's
// write-back).  Two cells with the same stamp hold the same positions, so a
// tissue whose cells all carry the stamps of its last write-back still has
// those positions on the device and skips the upload.
uint64_t next_stamp();

struct HandleDeleter {
  void operator()(dpm_tissue* h) const;
};
using Handle = std::shared_ptr<dpm_tissue>;

// Polygon cell, cell.hpp:11-25 / cell.cpp:12-60.
struct Cell2D {
  unsigned int NV = 0;
  float calA0 = 0, a0 = 0, l0 = 0, r0 = 0;
  float Ka = 0, Kb = 0, Kl = 0, Ks = 0;  // uninitialised in the reference; 0 here
  std::vector<std::array<float, 2>> Verticies;
  std::vector<std::array<float, 2>> Forces;
  uint64_t verts_stamp = next_stamp();
  void SetVerts(std::vector<std::array<float, 2>> v) {
    Verticies = std::move(v);
    verts_stamp = next_stamp();
  }
  // The fifth argument is the circumradius (cell.cpp:12-16), although the
  // reference header names it a0.
  Cell2D(float x0, float y0, float calA, unsigned int numVerts, float r0);
  float GetArea() const;
  float GetPerim() const;
};

// Icosphere cell, cell.hpp:28-58 / cell.cpp:62-288, generalised to any
// subdivision level (the reference fixes 2: 162 vertices, 320 faces).
struct Cell3D {
  unsigned int NV = 0, NF = 0;
  int subdiv = 2;
  float calA0 = 0, r0 = 0, v0 = 0, sa0 = 0, a0 = 0;
  float Kv = 0, Ka = 0, Ks = 0;  // Ks is uninitialised in the reference; 0 here
  float Kb = 0;                  // bending (no reference counterpart)
  float Volume = 0, SurfaceArea = 0;
  std::vector<std::array<float, 3>> Verts;
  std::vector<std::array<float, 3>> Forces;
  std::vector<std::array<unsigned int, 3>> Faces;
  uint64_t verts_stamp = next_stamp();
  void SetVerts(std::vector<std::array<float, 3>> v) {
    Verts = std::move(v);
    verts_stamp = next_stamp();
  }

  Cell3D(std::array<float, 3> starting_point, float calA, float r0, int subdiv = 2);
  float GetVolume() const;
  float GetSurfaceArea() const;
  std::vector<std::vector<float>> GetPositions() const;                // 3 x NV
  std::vector<std::vector<float>> GetVesselPositions(float L) const;   // 3 x NV
  std::vector<std::vector<float>> GetForces() const;                   // 3 x NV
  std::vector<std::array<int, 3>> GetFaces() const;                    // NF x 3
  std::array<float, 3> GetCOM() const;

 private:
  std::vector<std::pair<int, int>> midpointCache;  // (key, vertex), as cell.cpp:160-194
  unsigned int AddMiddlePoint(unsigned int p1, unsigned int p2);
};

// Tissue.hpp:35-55, Tissue2D.cpp.
struct Tissue2D {
  std::vector<Cell2D> cells;
  int NCELLS = 0;
  bool PBC = true;
  float Kre = 1.0f, Kat = 0.0f, phi0 = 0.0f, L = 0.0f;
  int maxNV = 0;

  Tissue2D(std::vector<Cell2D> cells, float packingFraction);
  void Disperse();
  void CLEulerUpdate(int nsteps, float dt);

 private:
  Handle h_;
  std::vector<int> h_nv_;
  std::vector<uint64_t> synced_;  // cell stamps of the positions the device holds
};

// Tissue.hpp:13-33, Tissue3D.cpp.
struct Tissue3D {
  int NCELLS = 0;
  int PBC = 1;
  float L = 0.0f;
  float Kre = 0.0f, Kat = 0.0f;  // uninitialised in the reference; 0 here
  std::vector<Cell3D> Cells;
  // Extensions (no reference counterpart): run the vertex-vertex adhesion
  // kernel the reference defines but never launches, and the substrate
  // stiffness gradient Ks*(1 + KsGradient*x/L).
  bool VertexAdhesion = false;
  float KsGradient = 0.0f;
  bool Verbose = true;  // print the reference's timing lines

  Tissue3D(std::vector<Cell3D> cells, float phi0);
  void CLEulerUpdate(int nsteps, float dt);
  void Disperse2D();

 private:
  Handle h_;
  int h_ncells_ = -1, h_nv_ = -1;
  std::vector<unsigned int> h_faces_;
  std::vector<uint64_t> synced_;  // cell stamps of the positions the device holds
};

// Throws std::invalid_argument for DPM_EINVAL, std::runtime_error otherwise.
void check(int status);

}  // namespace dpmhost
