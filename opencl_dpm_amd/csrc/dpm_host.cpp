// Host-side model behind `clDPM` (see dpm_host.hpp).  Constructor and
// dispersal arithmetic keeps the reference's float/double mix so that the
// generated geometry is bit-identical: float variables, double literals and
// double-returning library calls exactly where the reference has them
// (src/cell.cpp, src/Tissue2D.cpp:15-114, src/Tissue3D.cpp:16-116; with the
// reference's headers, sqrt/round of a float bind to the float overloads and
// cbrt/pow/cos/sin of a float argument with double operands compute in
// double).
#include "dpm_host.hpp"

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <stdexcept>

#include "dpm.h"

namespace dpmhost {

uint64_t next_stamp() {
  static std::atomic<uint64_t> counter{0};
  return ++counter;
}

void HandleDeleter::operator()(dpm_tissue* h) const { dpm_destroy(h); }

void check(int status) {
  if (status == DPM_OK) return;
  std::string msg = dpm_last_error();
  if (status == DPM_EINVAL) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

// Python may assign Verts and Forces lists of any length; the getters read
// NV entries (the reference indexes past the end of a shorter list).
static void require_entries(size_t have, unsigned int nv, const char* what) {
  if (have < nv)
    throw std::invalid_argument(std::string(what) + " has " + std::to_string(have) +
                                " entries, fewer than NV = " + std::to_string(nv));
}

// ------------------------------------------------------------------ Cell2D

Cell2D::Cell2D(float x0, float y0, float CalA, unsigned int numVerts, float _r0) {
  NV = numVerts;
  r0 = _r0;
  if (NV < 3) throw std::invalid_argument("Cell2D needs at least 3 vertices");
  calA0 = CalA * (NV * tan(M_PI / NV) / M_PI);
  Verticies.resize(NV);
  Forces.resize(NV);
  for (unsigned int i = 0; i < NV; i++) {
    Verticies[i][0] = r0 * (cos(2.0 * M_PI * (i + 1.0) / (float)NV)) + x0;
    Verticies[i][1] = r0 * (sin(2.0 * M_PI * (i + 1.0) / (float)NV)) + y0;
    Forces[i][0] = 0.0f;
    Forces[i][1] = 0.0f;
  }
  a0 = GetArea();
  l0 = 2.0 * sqrt(M_PI * calA0 * a0) / (float)NV;
}

float Cell2D::GetArea() const {
  require_entries(Verticies.size(), NV, "Verts");
  float area = 0.0f;
  unsigned int j = NV - 1;
  for (unsigned int i = 0; i < NV; i++) {
    area += 0.5 * ((Verticies[j][0] + Verticies[i][0]) * (Verticies[j][1] - Verticies[i][1]));
    j = i;
  }
  if (area < 0.0f) area = -area;
  return area;
}

// Closed-polygon perimeter.  (The reference's GetPerim, cell.cpp:49-60,
// reads one vertex past the end for its closing edge; this closes the
// polygon on vertex NV-1 -> 0.)
float Cell2D::GetPerim() const {
  require_entries(Verticies.size(), NV, "Verts");
  float dist = 0.0f;
  for (unsigned int i = 0; i < NV; i++) {
    const unsigned int j = (i + 1) % NV;
    const float dx = Verticies[j][0] - Verticies[i][0];
    const float dy = Verticies[j][1] - Verticies[i][1];
    dist += sqrtf(dx * dx + dy * dy);
  }
  return dist;
}

// ------------------------------------------------------------------ Cell3D

Cell3D::Cell3D(std::array<float, 3> starting_point, float calA, float _r0, int _subdiv) {
  if (_subdiv < 0 || _subdiv > 3)
    throw std::invalid_argument("icosphere subdivision must be 0..3");
  subdiv = _subdiv;
  calA0 = calA;
  r0 = _r0;
  Kv = 0.0f;
  Ka = 0.0f;
  Verts.resize(12);
  // Icosahedron: golden-ratio vertices, normalised (cell.cpp:69-93).
  const float t = (1 + std::sqrt(5.0)) / 2;
  const float base[12][3] = {{-1, t, 0}, {1, t, 0},  {-1, -t, 0}, {1, -t, 0},
                             {0, -1, t}, {0, 1, t},  {0, -1, -t}, {0, 1, -t},
                             {t, 0, -1}, {t, 0, 1},  {-t, 0, -1}, {-t, 0, 1}};
  for (int i = 0; i < 12; i++) {
    for (int k = 0; k < 3; k++) Verts[i][k] = base[i][k];
    const float norm = sqrtf(Verts[i][0] * Verts[i][0] + Verts[i][1] * Verts[i][1] +
                             Verts[i][2] * Verts[i][2]);
    for (int k = 0; k < 3; k++) Verts[i][k] /= norm;
  }
  // The 20 faces in the reference's order (cell.cpp:94-119).
  static const unsigned int F20[20][3] = {
      {0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
      {11, 10, 2}, {10, 7, 6}, {7, 1, 8},  {3, 9, 4},  {3, 4, 2},   {3, 2, 6}, {3, 6, 8},
      {3, 8, 9},  {4, 9, 5},  {2, 4, 11},  {6, 2, 10}, {8, 6, 7},   {9, 8, 1}};
  for (int f = 0; f < 20; f++) Faces.push_back({F20[f][0], F20[f][1], F20[f][2]});
  // Each pass splits every face in four (cell.cpp:121-142).  The midpoint
  // cache is shared by all passes, as in the reference.
  for (int pass = 0; pass < subdiv; pass++) {
    std::vector<std::array<unsigned int, 3>> next;
    next.reserve(Faces.size() * 4);
    for (size_t j = 0; j < Faces.size(); j++) {
      const unsigned int a = AddMiddlePoint(Faces[j][0], Faces[j][1]);
      const unsigned int b = AddMiddlePoint(Faces[j][1], Faces[j][2]);
      const unsigned int c = AddMiddlePoint(Faces[j][2], Faces[j][0]);
      next.push_back({Faces[j][0], a, c});
      next.push_back({Faces[j][1], b, a});
      next.push_back({Faces[j][2], c, b});
      next.push_back({a, b, c});
    }
    Faces.swap(next);
  }
  NV = (unsigned int)Verts.size();
  NF = (unsigned int)Faces.size();
  Forces.assign(NV, {0.0f, 0.0f, 0.0f});
  for (unsigned int vi = 0; vi < NV; vi++)
    for (int k = 0; k < 3; k++) {
      Verts[vi][k] *= r0;
      Verts[vi][k] += starting_point[k];
    }
  // Sphere targets (cell.cpp:153-155).
  v0 = (4.0f / 3.0f) * M_PI * std::pow((double)r0, 3);
  sa0 = std::pow((6 * std::sqrt(M_PI) * v0 * calA), (double)(2.0f / 3.0f));
  a0 = (sa0 / (float)NF);
  Volume = GetVolume();
  SurfaceArea = GetSurfaceArea();
}

// Midpoint of edge (p1,p2), projected to the unit sphere with the reference's
// component-by-component normalisation (cell.cpp:160-194): the norm is
// recomputed after each component is divided.
unsigned int Cell3D::AddMiddlePoint(unsigned int p1, unsigned int p2) {
  const unsigned int s = p1 + p2;
  const int key = (int)(s * (s + 1) / 2 + (p1 < p2 ? p1 : p2));
  for (const auto& e : midpointCache)
    if (e.first == key) return (unsigned int)e.second;
  const std::array<float, 3> v1 = Verts[p2], v2 = Verts[p1];
  std::array<float, 3> m;
  for (int i = 0; i < 3; i++) {
    m[i] = v1[i] + v2[i];
    m[i] *= 0.5;
  }
  for (int i = 0; i < 3; i++) {
    const float norm = sqrtf(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    m[i] /= norm;
  }
  Verts.push_back(m);
  const int idx = (int)Verts.size() - 1;
  midpointCache.emplace_back(key, idx);
  return (unsigned int)idx;
}

float Cell3D::GetVolume() const {
  require_entries(Verts.size(), NV, "Verts");
  float volume = 0.0f;
  for (const auto& tri : Faces) {
    const auto& p0 = Verts[tri[0]];
    const auto& p1 = Verts[tri[1]];
    const auto& p2 = Verts[tri[2]];
    const float c0 = p1[1] * p2[2] - p1[2] * p2[1];
    const float c1 = p1[2] * p2[0] - p1[0] * p2[2];
    const float c2 = p1[0] * p2[1] - p1[1] * p2[0];
    const float part = c0 * p0[0] + c1 * p0[1] + c2 * p0[2];
    volume += part;
  }
  return std::abs(volume) / 6.0;
}

// Sum of |cross| over faces: twice the triangulated area, as the reference
// returns it (cell.cpp:213-230).
float Cell3D::GetSurfaceArea() const {
  require_entries(Verts.size(), NV, "Verts");
  float sa = 0.0f;
  for (const auto& tri : Faces) {
    const auto& p0 = Verts[tri[0]];
    const auto& p1 = Verts[tri[1]];
    const auto& p2 = Verts[tri[2]];
    const float A0 = p1[0] - p0[0], A1 = p1[1] - p0[1], A2 = p1[2] - p0[2];
    const float B0 = p2[0] - p0[0], B1 = p2[1] - p0[1], B2 = p2[2] - p0[2];
    const float c0 = A1 * B2 - A2 * B1;
    const float c1 = A2 * B0 - A0 * B2;
    const float c2 = A0 * B1 - A1 * B0;
    sa += sqrtf(c0 * c0 + c1 * c1 + c2 * c2);
  }
  return sa;
}

std::vector<std::vector<float>> Cell3D::GetPositions() const {
  require_entries(Verts.size(), NV, "Verts");
  std::vector<std::vector<float>> p(3, std::vector<float>(NV));
  for (unsigned int i = 0; i < NV; i++)
    for (int k = 0; k < 3; k++) p[k][i] = Verts[i][k];
  return p;
}

// Positions wrapped on a cylinder of circumference L (cell.cpp:242-253).
std::vector<std::vector<float>> Cell3D::GetVesselPositions(float L) const {
  require_entries(Verts.size(), NV, "Verts");
  std::vector<std::vector<float>> p(3, std::vector<float>(NV));
  const float scale = (2.0 * M_PI) / L;
  const float radius = L / (2 * M_PI);
  for (unsigned int vi = 0; vi < NV; vi++) {
    const float theta = Verts[vi][0] * scale;
    p[0][vi] = (radius - Verts[vi][2]) * cosf(theta);
    p[2][vi] = (radius - Verts[vi][2]) * sinf(theta);
    p[1][vi] = Verts[vi][1];
  }
  return p;
}

std::vector<std::vector<float>> Cell3D::GetForces() const {
  require_entries(Forces.size(), NV, "Forces");
  std::vector<std::vector<float>> p(3, std::vector<float>(NV));
  for (unsigned int i = 0; i < NV; i++)
    for (int k = 0; k < 3; k++) p[k][i] = Forces[i][k];
  return p;
}

std::vector<std::array<int, 3>> Cell3D::GetFaces() const {
  std::vector<std::array<int, 3>> f(NF);
  for (unsigned int i = 0; i < NF; i++)
    for (int k = 0; k < 3; k++) f[i][k] = (int)Faces[i][k];
  return f;
}

std::array<float, 3> Cell3D::GetCOM() const {
  require_entries(Verts.size(), NV, "Verts");
  std::array<float, 3> com = {0.0f, 0.0f, 0.0f};
  for (unsigned int vi = 0; vi < NV; vi++)
    for (int k = 0; k < 3; k++) com[k] += Verts[vi][k];
  for (int k = 0; k < 3; k++) com[k] /= (float)NV;
  return com;
}

// ---------------------------------------------------------------- Tissue2D

Tissue2D::Tissue2D(std::vector<Cell2D> inputCells, float packingFraction) {
  cells = std::move(inputCells);
  NCELLS = (int)cells.size();
  if (NCELLS == 0) throw std::invalid_argument("Tissue2D needs at least one cell");
  PBC = true;
  phi0 = packingFraction;
  float totalArea = 0.0f;
  maxNV = (int)cells[0].NV;
  for (const auto& c : cells) {
    totalArea += c.GetArea();
    if ((int)c.NV > maxNV) maxNV = (int)c.NV;
  }
  L = sqrtf(totalArea) / phi0;
  Kre = 1.0f;
  Kat = 0.0f;
}

namespace {
// Dispersal of Tissue2D.cpp:42-105 and Tissue3D.cpp:40-106: starting centres
// drawn from the process-wide drand48 stream (x then y per cell, as the
// reference draws them), then the soft-disk relaxation with contact
// diameter scale * (r_i + r_j) on the device (dpm_disperse, bit-identical
// to the reference's host loop).
void relax_centres(const std::vector<float>& radius, float scale, float L, std::vector<float>& X,
                   std::vector<float>& Y, bool warn) {
  const int n = (int)radius.size();
  X.assign(n, 0.0f);
  Y.assign(n, 0.0f);
  std::vector<float> r(n);
  for (int i = 0; i < n; i++) {
    X[i] = drand48() * L;
    Y[i] = drand48() * L;
    r[i] = radius[i] * scale;
  }
  int res[2] = {0, 0};
  check(dpm_disperse(r.data(), n, L, X.data(), Y.data(), 100000, res));
  if (res[1] && warn) std::cerr << "Warning: Max timesteps for dispersion reached" << std::endl;
}
}  // namespace

void Tissue2D::Disperse() {
  NCELLS = (int)cells.size();  // Cells may have been reassigned from Python
  for (const auto& c : cells) require_entries(c.Verticies.size(), c.NV, "Verts");
  std::vector<float> r(NCELLS), X, Y;
  for (int i = 0; i < NCELLS; i++) r[i] = cells[i].r0;
  relax_centres(r, 1.0f, L, X, Y, false);
  for (int i = 0; i < NCELLS; i++) {
    for (unsigned int j = 0; j < cells[i].NV; j++) {
      cells[i].Verticies[j][0] = cells[i].r0 * (cos(2.0 * M_PI * (j + 1) / cells[i].NV)) + X[i];
      cells[i].Verticies[j][1] = cells[i].r0 * (sin(2.0 * M_PI * (j + 1) / cells[i].NV)) + Y[i];
    }
    cells[i].verts_stamp = next_stamp();
  }
}

void Tissue2D::CLEulerUpdate(int nsteps, float dt) {
  // Tissue2D.cpp:116-241.  With nsteps <= 0 the reference runs no step and
  // hands back its zero-initialised force buffer.
  if (nsteps <= 0) {
    for (auto& c : cells)
      for (auto& f : c.Forces) f = {0.0f, 0.0f};
    return;
  }
  const int C = NCELLS = (int)cells.size();
  std::vector<int> nv(C);
  std::vector<float> Ka(C), Kl(C), Kb(C), a0(C), l0(C), r0(C);
  size_t V = 0;
  for (int c = 0; c < C; c++) {
    // The reference packs the first NV entries of Verticies and writes back
    // NV entries of Verticies and Forces (Tissue2D.cpp:137-139, 235-238);
    // a Verts list assigned shorter than NV from Python is an error here
    // (the reference reads past its end), a longer one keeps its tail.
    if (cells[c].Verticies.size() < cells[c].NV)
      throw std::invalid_argument("cell " + std::to_string(c) + ": Verts has " +
                                  std::to_string(cells[c].Verticies.size()) +
                                  " entries, fewer than NV = " + std::to_string(cells[c].NV));
    nv[c] = (int)cells[c].NV;
    Ka[c] = cells[c].Ka;
    Kl[c] = cells[c].Kl;
    Kb[c] = cells[c].Kb;
    a0[c] = cells[c].a0;
    l0[c] = cells[c].l0;
    r0[c] = cells[c].r0;
    V += nv[c];
  }
  const int cb[2] = {0, C};
  const int pbc = PBC ? 1 : 0;
  bool resident = (bool)h_ && h_nv_ == nv && (int)synced_.size() == C;
  for (int c = 0; c < C && resident; c++) resident = cells[c].verts_stamp == synced_[c];
  if (!h_ || h_nv_ != nv) {
    dpm2d_desc d{};
    d.ntissues = 1;
    d.cell_begin = cb;
    d.nv = nv.data();
    d.Ka = Ka.data();
    d.Kl = Kl.data();
    d.Kb = Kb.data();
    d.a0 = a0.data();
    d.l0 = l0.data();
    d.r0 = r0.data();
    d.L = &L;
    d.pbc = &pbc;
    d.Kre = &Kre;
    d.Kat = &Kat;
    d.mask = DPM2D_ALL;
    dpm_tissue* raw = nullptr;
    h_.reset();
    check(dpm2d_create(&d, &raw));
    h_ = Handle(raw, HandleDeleter());
    h_nv_ = nv;
  } else {
    check(dpm2d_set_cell_params(h_.get(), Ka.data(), Kl.data(), Kb.data(), a0.data(), l0.data(),
                                r0.data()));
    check(dpm2d_set_tissue_params(h_.get(), &L, &pbc, &Kre, &Kat, DPM2D_ALL));
  }
  std::vector<float> xy(2 * V);
  size_t k = 0;
  if (!resident) {  // the device holds other positions: upload these
    synced_.clear();
    for (const auto& c : cells)
      for (unsigned int vi = 0; vi < c.NV; vi++) {
        xy[k++] = c.Verticies[vi][0];
        xy[k++] = c.Verticies[vi][1];
      }
    check(dpm2d_set_positions(h_.get(), xy.data(), V));
  }
  check(dpm_euler_async(h_.get(), nsteps, dt));
  check(dpm_synchronize(h_.get()));
  std::vector<float> f(2 * V);
  check(dpm2d_get_positions(h_.get(), xy.data(), V));
  check(dpm2d_get_forces(h_.get(), f.data(), V));
  k = 0;
  synced_.resize(C);
  for (int ci = 0; ci < C; ci++) {
    Cell2D& c = cells[ci];
    if (c.Forces.size() < c.NV) c.Forces.resize(c.NV);
    for (unsigned int vi = 0; vi < c.NV; vi++, k++) {
      c.Verticies[vi] = {xy[2 * k], xy[2 * k + 1]};
      c.Forces[vi] = {f[2 * k], f[2 * k + 1]};
    }
    c.verts_stamp = synced_[ci] = next_stamp();
  }
}

// ---------------------------------------------------------------- Tissue3D

Tissue3D::Tissue3D(std::vector<Cell3D> cells, float phi0) {
  Cells = std::move(cells);
  NCELLS = (int)Cells.size();
  float volume = 0.0f;
  for (const auto& c : Cells) volume += c.v0;
  L = (float)(std::cbrt((double)volume) / phi0);
  PBC = 1;
}

void Tissue3D::Disperse2D() {
  NCELLS = (int)Cells.size();  // Cells may have been reassigned from Python
  for (const auto& c : Cells) require_entries(c.Verts.size(), c.NV, "Verts");
  std::vector<float> r(NCELLS), X, Y;
  for (int i = 0; i < NCELLS; i++) r[i] = Cells[i].r0;
  relax_centres(r, 2.0f, L, X, Y, true);
  for (int i = 0; i < NCELLS; i++) {
    const std::array<float, 3> com = Cells[i].GetCOM();
    for (unsigned int j = 0; j < Cells[i].NV; j++) {
      Cells[i].Verts[j][0] -= com[0];
      Cells[i].Verts[j][1] -= com[1];
      Cells[i].Verts[j][0] -= X[i];
      Cells[i].Verts[j][1] -= Y[i];
    }
    Cells[i].verts_stamp = next_stamp();
  }
}

void Tissue3D::CLEulerUpdate(int nsteps, float dt) {
  // Input validation as Tissue3D.cpp:123-197.
  if (nsteps <= 0) {
    std::cerr << "[ERROR] Invalid nsteps: " << nsteps << std::endl;
    throw std::invalid_argument("nsteps must be positive");
  }
  if (dt <= 0.0f || dt > 0.1f) {
    std::cerr << "[ERROR] Invalid dt: " << dt << " (should be in range (0, 0.1])" << std::endl;
    throw std::invalid_argument("dt must be positive and reasonable");
  }
  NCELLS = (int)Cells.size();
  if (NCELLS <= 0) {
    std::cerr << "[ERROR] Invalid NCELLS: " << NCELLS << std::endl;
    throw std::invalid_argument("NCELLS must be positive");
  }
  const Cell3D& c0 = Cells[0];
  const int nv = (int)c0.NV, nf = (int)c0.NF;
  std::vector<int> faces(3 * nf);
  std::vector<unsigned int> faces_u(3 * nf);
  for (int f = 0; f < nf; f++) {
    for (int k = 0; k < 3; k++) {
      faces[3 * f + k] = (int)c0.Faces[f][k];
      faces_u[3 * f + k] = c0.Faces[f][k];
    }
    if (c0.Faces[f][0] >= (unsigned)nv || c0.Faces[f][1] >= (unsigned)nv ||
        c0.Faces[f][2] >= (unsigned)nv) {
      std::cerr << "[ERROR] Invalid face index at face " << f << std::endl;
      throw std::runtime_error("Invalid face indices");
    }
  }
  const int C = NCELLS;
  std::vector<float> Kv(C), Ka(C), Ks(C), Kb(C), v0(C), a0(C);
  bool bending = false;
  std::vector<float> xyz((size_t)C * nv * 3);
  // The device still holds every cell's positions when each cell carries the
  // stamp of the last write-back (they were checked finite after that call).
  bool resident = (bool)h_ && h_ncells_ == C && h_nv_ == nv && h_faces_ == faces_u && (int)synced_.size() == C;
  for (int c = 0; c < C && resident; c++) resident = Cells[c].verts_stamp == synced_[c];
  for (int c = 0; c < C; c++) {
    const Cell3D& cell = Cells[c];
    if (cell.Kv <= 0 || cell.Ka <= 0 || cell.Ks <= 0) {
      std::cerr << "[ERROR] Invalid spring constants for cell " << c << ": Kv=" << cell.Kv
                << ", Ka=" << cell.Ka << ", Ks=" << cell.Ks << std::endl;
      throw std::runtime_error("Invalid spring constants");
    }
    if (cell.v0 <= 0 || cell.a0 <= 0) {
      std::cerr << "[ERROR] Invalid reference values for cell " << c << ": v0=" << cell.v0
                << ", a0=" << cell.a0 << std::endl;
      throw std::runtime_error("Invalid reference values");
    }
    if ((int)cell.Verts.size() != nv) throw std::runtime_error("all cells must share one mesh");
    Kv[c] = cell.Kv;
    Ka[c] = cell.Ka;
    Ks[c] = cell.Ks;
    Kb[c] = cell.Kb;
    bending = bending || cell.Kb != 0.0f;
    v0[c] = cell.v0;
    a0[c] = cell.a0;
    for (int v = 0; v < nv && !resident; v++)
      for (int k = 0; k < 3; k++) {
        const float val = cell.Verts[v][k];
        if (!std::isfinite(val)) {
          std::cerr << "[ERROR] Non-finite vertex coordinate at cell " << c << ", vertex " << v
                    << ", coord " << k << ": " << val << std::endl;
          throw std::runtime_error("Non-finite vertex coordinates");
        }
        xyz[((size_t)c * nv + v) * 3 + k] = val;
      }
  }
  unsigned mask = DPM3D_REFERENCE;
  if (bending) mask |= DPM3D_BENDING;
  if (VertexAdhesion) mask |= DPM3D_ADHESION;
  const int cb[2] = {0, C};
  const int pbc = PBC ? 1 : 0;
  if (!h_ || h_ncells_ != C || h_nv_ != nv || h_faces_ != faces_u) {
    dpm3d_desc d{};
    d.ntissues = 1;
    d.cell_begin = cb;
    d.nv = nv;
    d.nf = nf;
    d.faces = faces.data();
    d.Kv = Kv.data();
    d.Ka = Ka.data();
    d.Ks = Ks.data();
    d.Kb = Kb.data();
    d.v0 = v0.data();
    d.a0 = a0.data();
    d.L = &L;
    d.pbc = &pbc;
    d.Kre = &Kre;
    d.Kat = &Kat;
    d.ks_grad = &KsGradient;
    d.mask = mask;
    dpm_tissue* raw = nullptr;
    h_.reset();
    check(dpm3d_create(&d, &raw));
    h_ = Handle(raw, HandleDeleter());
    h_ncells_ = C;
    h_nv_ = nv;
    h_faces_ = faces_u;
  } else {
    check(dpm3d_set_cell_params(h_.get(), Kv.data(), Ka.data(), Ks.data(), Kb.data(), v0.data(),
                                a0.data()));
    check(dpm3d_set_tissue_params(h_.get(), &L, &pbc, &Kre, &Kat, &KsGradient, mask));
  }
  if (!resident) {
    synced_.clear();
    check(dpm3d_set_positions(h_.get(), xyz.data(), (size_t)C * nv));
  }
  const auto t0 = std::chrono::high_resolution_clock::now();
  check(dpm_euler_async(h_.get(), nsteps, dt));
  check(dpm_synchronize(h_.get()));
  const auto t1 = std::chrono::high_resolution_clock::now();
  if (Verbose) {
    const auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0);
    std::cout << nsteps << " timesteps completed in " << ms.count() << " ms" << std::endl;
    std::cout << "Average time per step: " << (ms.count() / static_cast<double>(nsteps)) << " ms"
              << std::endl;
  }
  std::vector<float> f((size_t)C * nv * 3), vol(C), sa(C);
  check(dpm3d_get_positions(h_.get(), xyz.data(), (size_t)C * nv));
  check(dpm3d_get_forces(h_.get(), f.data(), (size_t)C * nv));
  // GetVolume / GetSurfaceArea of every cell (Tissue3D.cpp:503-504), on the
  // device in the host functions' float order (bit-identical)
  check(dpm3d_cell_measures(h_.get(), vol.data(), sa.data(), nullptr));
  // Result checks and host reductions as Tissue3D.cpp:477-521.
  for (int c = 0; c < C; c++) {
    Cell3D& cell = Cells[c];
    if ((int)cell.Forces.size() < nv) cell.Forces.resize(nv);
    for (int v = 0; v < nv; v++) {
      for (int k = 0; k < 3; k++) {
        const size_t i = ((size_t)c * nv + v) * 3 + k;
        if (!std::isfinite(xyz[i])) {
          std::cerr << "[ERROR] Non-finite result vertex at cell " << c << ", vertex " << v
                    << ", coord " << k << ": " << xyz[i] << std::endl;
          throw std::runtime_error("Non-finite simulation results");
        }
        if (!std::isfinite(f[i]))
          std::cerr << "[WARNING] Non-finite force at cell " << c << ", vertex " << v << ", coord "
                    << k << ": " << f[i] << std::endl;
      }
      const size_t i = ((size_t)c * nv + v) * 3;
      cell.Verts[v] = {xyz[i], xyz[i + 1], xyz[i + 2]};
      cell.Forces[v] = {f[i], f[i + 1], f[i + 2]};
    }
    cell.verts_stamp = next_stamp();
    cell.Volume = vol[c];
    cell.SurfaceArea = sa[c];
    if (!std::isfinite(cell.Volume) || cell.Volume <= 0)
      std::cerr << "[WARNING] Invalid volume for cell " << c << ": " << cell.Volume << std::endl;
    if (!std::isfinite(cell.SurfaceArea) || cell.SurfaceArea <= 0)
      std::cerr << "[WARNING] Invalid surface area for cell " << c << ": " << cell.SurfaceArea
                << std::endl;
  }
  synced_.resize(C);
  for (int c = 0; c < C; c++) synced_[c] = Cells[c].verts_stamp;
}

}  // namespace dpmhost
