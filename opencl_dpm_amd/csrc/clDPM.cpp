// Python module `clDPM`: the reference's Python surface (src/DPMWrapper.cpp:11-16,
// src/CellWrapper.cpp:7-30, src/TissueWrapper.cpp:7-30) bound to the host
// model of dpm_host.hpp, whose timestep runs on libdpm_hip.  Attribute
// names, argument order, return shapes and exception types are the
// reference's; names marked "extension" have no reference counterpart.
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstdlib>

#include "dpm.h"
#include "dpm_host.hpp"

namespace py = pybind11;
using namespace dpmhost;

PYBIND11_MODULE(clDPM, m) {
  m.doc() = "Deformable Particle Model timestep on MI355X (HIP, gfx950)";

  py::class_<Cell2D>(m, "Cell2D")
      .def(py::init<float, float, float, unsigned int, float>())
      .def_readwrite("Ka", &Cell2D::Ka)
      .def_readwrite("Kl", &Cell2D::Kl)
      .def_readwrite("Kb", &Cell2D::Kb)
      .def_property("Verts", [](const Cell2D& c) { return c.Verticies; }, &Cell2D::SetVerts)
      .def_readwrite("Forces", &Cell2D::Forces)
      // extensions
      .def_readonly("NV", &Cell2D::NV)
      .def_readonly("a0", &Cell2D::a0)
      .def_readonly("l0", &Cell2D::l0)
      .def_readonly("r0", &Cell2D::r0)
      .def_readonly("calA0", &Cell2D::calA0)
      .def("GetArea", &Cell2D::GetArea)
      .def("GetPerim", &Cell2D::GetPerim);

  py::class_<Tissue2D>(m, "Tissue2D")
      .def(py::init<std::vector<Cell2D>, float>())
      .def_readwrite("Cells", &Tissue2D::cells)
      .def_readonly("NCELLS", &Tissue2D::NCELLS)
      .def_readonly("L", &Tissue2D::L)
      .def_readonly("PBC", &Tissue2D::PBC)
      .def_readwrite("Kre", &Tissue2D::Kre)
      .def_readwrite("Kat", &Tissue2D::Kat)
      .def("CLEulerUpdate", &Tissue2D::CLEulerUpdate, py::call_guard<py::gil_scoped_release>())
      .def("Disperse", &Tissue2D::Disperse);

  py::class_<Cell3D>(m, "Cell3D")
      .def(py::init<std::array<float, 3>, float, float, int>(), py::arg("starting_point"),
           py::arg("calA"), py::arg("r0"), py::arg("subdiv") = 2)
      .def_readwrite("Kv", &Cell3D::Kv)
      .def_readwrite("Ka", &Cell3D::Ka)
      .def_readwrite("Ks", &Cell3D::Ks)
      .def_property("Verts", [](const Cell3D& c) { return c.Verts; }, &Cell3D::SetVerts)
      .def("GetVolume", &Cell3D::GetVolume)
      .def("GetPositions", &Cell3D::GetPositions)
      .def("GetVesselPositions", &Cell3D::GetVesselPositions)
      .def("GetFaces", &Cell3D::GetFaces)
      .def("GetForces", &Cell3D::GetForces)
      // extensions
      .def_readwrite("Kb", &Cell3D::Kb)
      .def_readwrite("Forces", &Cell3D::Forces)
      .def_readonly("NV", &Cell3D::NV)
      .def_readonly("NF", &Cell3D::NF)
      .def_readonly("v0", &Cell3D::v0)
      .def_readonly("sa0", &Cell3D::sa0)
      .def_readonly("a0", &Cell3D::a0)
      .def_readonly("r0", &Cell3D::r0)
      .def_readonly("calA0", &Cell3D::calA0)
      .def_readonly("Volume", &Cell3D::Volume)
      .def_readonly("SurfaceArea", &Cell3D::SurfaceArea)
      .def("GetSurfaceArea", &Cell3D::GetSurfaceArea)
      .def("GetCOM", &Cell3D::GetCOM);

  py::class_<Tissue3D>(m, "Tissue3D")
      .def(py::init<std::vector<Cell3D>, float>())
      .def_readwrite("Kre", &Tissue3D::Kre)
      .def_readwrite("Kat", &Tissue3D::Kat)
      .def_readwrite("Cells", &Tissue3D::Cells)
      .def_readonly("NCELLS", &Tissue3D::NCELLS)
      .def_readonly("L", &Tissue3D::L)
      .def_readonly("PBC", &Tissue3D::PBC)
      .def("CLEulerUpdate", &Tissue3D::CLEulerUpdate, py::call_guard<py::gil_scoped_release>())
      .def("Disperse2D", &Tissue3D::Disperse2D)
      // extensions
      .def_readwrite("VertexAdhesion", &Tissue3D::VertexAdhesion)
      .def_readwrite("KsGradient", &Tissue3D::KsGradient)
      .def_readwrite("Verbose", &Tissue3D::Verbose);

  // extensions: the reference draws its dispersal from the process-wide
  // drand48 stream; these set that stream's state.
  m.def("seed48", [](unsigned short a, unsigned short b, unsigned short c) {
    unsigned short s[3] = {a, b, c};
    seed48(s);
  });
  m.def("srand48", [](long s) { srand48(s); });
  m.def("device_count", []() {
    int n = 0;
    dpm_device_count(&n);
    return n;
  });
}
