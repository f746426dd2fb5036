// TEST INFRASTRUCTURE ONLY — a C-ABI shim over the reference's own host C++
// (src/cell.cpp, src/Tissue2D.cpp, src/Tissue3D.cpp compiled from where they
// lie under the reference checkout by oracle/Makefile into oracle/_ref/).
// It exposes only the host-side set-up paths that need no OpenCL device:
// cell construction (cell.cpp:12-33, 62-194), the host reductions
// (cell.cpp:35-47, 196-230, 276-288) and the dispersal routines
// (Tissue2D.cpp:42-114, Tissue3D.cpp:40-116).  Used to write golden fixtures
// and to cross-check the product's constructors; never shipped.
#include "Tissue.hpp"
#include "cell.hpp"
#include <cstdlib>
#include <cstring>
#include <unistd.h>
#include <vector>

extern "C" {

int refshim_version(void) { return 1; }

void refshim_seed48(unsigned short s0, unsigned short s1, unsigned short s2) {
  unsigned short s[3] = {s0, s1, s2};
  seed48(s);
}

// Cell2D(x0, y0, calA, NV, r0): vertices (NV*2), then a0, l0, calA0, area.
void refshim_cell2d(float x0, float y0, float calA, unsigned nv, float r0,
                    float *verts, float *scalars) {
  DPM::Cell2D c(x0, y0, calA, nv, r0);
  for (unsigned i = 0; i < nv; i++) {
    verts[2 * i] = c.Verticies[i][0];
    verts[2 * i + 1] = c.Verticies[i][1];
  }
  scalars[0] = c.a0;
  scalars[1] = c.l0;
  scalars[2] = c.calA0;
  scalars[3] = c.GetArea();
}

// Cell2D::GetArea on caller-given vertices.
float refshim_cell2d_area(unsigned nv, const float *verts) {
  DPM::Cell2D c(0.0f, 0.0f, 1.0f, nv, 1.0f);
  for (unsigned i = 0; i < nv; i++) {
    c.Verticies[i][0] = verts[2 * i];
    c.Verticies[i][1] = verts[2 * i + 1];
  }
  return c.GetArea();
}

// Cell3D(sp, calA, r0): vertices (162*3), faces (320*3), then
// v0, sa0, a0, Volume, SurfaceArea, calA0, r0.
void refshim_cell3d(const float *sp, float calA, float r0, float *verts,
                    int *faces, float *scalars) {
  DPM::Cell3D c({sp[0], sp[1], sp[2]}, calA, r0);
  for (unsigned i = 0; i < DPM::Cell3D::NV; i++)
    for (int k = 0; k < 3; k++) verts[3 * i + k] = c.Verts[i][k];
  for (unsigned f = 0; f < DPM::Cell3D::NF; f++)
    for (int k = 0; k < 3; k++) faces[3 * f + k] = (int)c.Faces[f][k];
  scalars[0] = c.v0;
  scalars[1] = c.sa0;
  scalars[2] = c.a0;
  scalars[3] = c.Volume;
  scalars[4] = c.SurfaceArea;
  scalars[5] = c.calA0;
  scalars[6] = c.r0;
}

// GetVolume, GetSurfaceArea, GetCOM on caller-given vertices (162*3).
void refshim_cell3d_measure(const float *verts, float *out5) {
  DPM::Cell3D c({0.0f, 0.0f, 0.0f}, 1.0f, 1.0f);
  for (unsigned i = 0; i < DPM::Cell3D::NV; i++)
    for (int k = 0; k < 3; k++) c.Verts[i][k] = verts[3 * i + k];
  out5[0] = c.GetVolume();
  out5[1] = c.GetSurfaceArea();
  std::array<float, 3> com = c.GetCOM();
  out5[2] = com[0];
  out5[3] = com[1];
  out5[4] = com[2];
}

// Tissue2D(cells, phi0) + Disperse().  Cells are given by (calA, NV, r0)
// triples, constructed at the origin as the reference callers do.  The
// constructor reads its kernel text from a cwd-relative path, so the caller
// passes the reference checkout as refdir.  Outputs: L, then vertices of all
// cells packed back to back (sum NV * 2).
int refshim_tissue2d_disperse(int ncells, const float *calA, const unsigned *nv,
                              const float *r0, float phi0, const char *refdir,
                              float *L, float *verts) {
  char cwd[4096];
  if (!getcwd(cwd, sizeof(cwd))) return -1;
  if (chdir(refdir) != 0) return -2;
  std::vector<DPM::Cell2D> cells;
  for (int i = 0; i < ncells; i++)
    cells.emplace_back(0.0f, 0.0f, calA[i], nv[i], r0[i]);
  DPM::Tissue2D T(cells, phi0);
  int rc = chdir(cwd);
  (void)rc;
  T.Disperse();
  *L = T.L;
  size_t k = 0;
  for (int i = 0; i < ncells; i++)
    for (unsigned v = 0; v < T.cells[i].NV; v++) {
      verts[k++] = T.cells[i].Verticies[v][0];
      verts[k++] = T.cells[i].Verticies[v][1];
    }
  return 0;
}

// Tissue3D(cells, phi0) + Disperse2D().  Cells given by (calA, r0) at the
// origin.  Outputs: L, then 162*3 vertices per cell.
int refshim_tissue3d_disperse2d(int ncells, const float *calA, const float *r0,
                                float phi0, const char *refdir, float *L,
                                float *verts) {
  char cwd[4096];
  if (!getcwd(cwd, sizeof(cwd))) return -1;
  if (chdir(refdir) != 0) return -2;
  std::vector<DPM::Cell3D> cells;
  for (int i = 0; i < ncells; i++)
    cells.emplace_back(std::array<float, 3>{0.0f, 0.0f, 0.0f}, calA[i], r0[i]);
  DPM::Tissue3D T(cells, phi0);
  int rc = chdir(cwd);
  (void)rc;
  T.Disperse2D();
  *L = T.L;
  size_t k = 0;
  for (int i = 0; i < ncells; i++)
    for (unsigned v = 0; v < DPM::Cell3D::NV; v++)
      for (int c = 0; c < 3; c++) verts[k++] = T.Cells[i].Verts[v][c];
  return 0;
}

}  // extern "C"
