// A small pybind11 module built the way the reference builds _contouring_cpp_module
// (src/cpp/bindings.cpp:73): the reference's grid, mesh and stats types registered as there
// (bindings.cpp:126-185, a subset of the stats fields), then the drop-in's one call,
// bind_hip_backend(m). Built and imported by tests/test_drop_in.py against the reference's
// include/dc_backend.h.
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "dc_backend.h"
#include "dc_hip_pybind.h"

namespace py = pybind11;

PYBIND11_MODULE(_dc_hip_dropin, m) {
  py::class_<DenseSdfGrid>(m, "DenseSdfGrid")
      .def(py::init<>())
      .def_readwrite("nx", &DenseSdfGrid::nx)
      .def_readwrite("ny", &DenseSdfGrid::ny)
      .def_readwrite("nz", &DenseSdfGrid::nz)
      .def_readwrite("vx", &DenseSdfGrid::vx)
      .def_readwrite("vy", &DenseSdfGrid::vy)
      .def_readwrite("vz", &DenseSdfGrid::vz)
      .def_readwrite("ox", &DenseSdfGrid::ox)
      .def_readwrite("oy", &DenseSdfGrid::oy)
      .def_readwrite("oz", &DenseSdfGrid::oz)
      .def_readwrite("values", &DenseSdfGrid::values);
  py::class_<DualContouringMesh>(m, "DualContouringMesh")
      .def(py::init<>())
      .def_readwrite("vertices", &DualContouringMesh::vertices)
      .def_readwrite("faces", &DualContouringMesh::faces);
  py::class_<DualContouringStats>(m, "DualContouringStats")
      .def(py::init<>())
      .def_readwrite("backend", &DualContouringStats::backend)
      .def_readwrite("vertex_count", &DualContouringStats::vertex_count)
      .def_readwrite("face_count", &DualContouringStats::face_count)
      .def_readwrite("total_ms", &DualContouringStats::total_ms);
  bind_hip_backend(m);
}
