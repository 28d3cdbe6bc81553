// dc_hip_pybind.h — pybind11 registration of HipDualContouringBackend for the reference's Python
// module (_contouring_cpp_module, src/cpp/bindings.cpp:73). Inside PYBIND11_MODULE, after the
// reference has registered DenseSdfGrid, DualContouringMesh and DualContouringStats
// (bindings.cpp:126-185), one call adds the class:
//
//   #include "dc_hip_pybind.h"
//   ...
//   bind_hip_backend(m);
//
// HipSparseDualContouringBackend and HipSparseMvdcDualContouringBackend mirror the reference's
// sparse pair (bindings.cpp:229-262). The dense class mirrors the reference's
// CudaDualContouringBackend binding (bindings.cpp:216-228):
// extract(grid) and extract_device(grid) return (mesh, stats); postprocess(mesh, close_holes,
// remove_floaters) is the reference's postprocess_mesh (bindings.cpp:145-146) on the GPU, and
// postprocess_last the same on the last extraction's mesh in device memory. The module functions
// mesh_to_sdf_hip / mesh_to_sdf_device_hip mirror mesh_to_sdf_cuda / mesh_to_sdf_device_cuda
// (bindings.cpp:270-284): (vertices (n, 3) float32, faces (m, 3) int32, nx, ny, nz, ox, oy, oz,
// vx, vy, vz) -> (grid, ms). DenseSdfGridDevice is registered
// here only when the build has no CUDA backends (with DC_ENABLE_CUDA the reference registers it,
// bindings.cpp:195-207); this registration adds d_values as an integer device address, so a
// torch tensor's data_ptr() can be passed.
#pragma once

#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstdint>
#include <string>

#include "dc_hip_backend.h"

inline void bind_hip_backend(pybind11::module_& m) {
  namespace py = pybind11;
#ifndef DC_ENABLE_CUDA
  py::class_<DenseSdfGridDevice>(m, "DenseSdfGridDevice")
      .def(py::init([]() { return DenseSdfGridDevice{0, 0, 0, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f,
                                                     nullptr}; }))
      .def_readwrite("nx", &DenseSdfGridDevice::nx)
      .def_readwrite("ny", &DenseSdfGridDevice::ny)
      .def_readwrite("nz", &DenseSdfGridDevice::nz)
      .def_readwrite("vx", &DenseSdfGridDevice::vx)
      .def_readwrite("vy", &DenseSdfGridDevice::vy)
      .def_readwrite("vz", &DenseSdfGridDevice::vz)
      .def_readwrite("ox", &DenseSdfGridDevice::ox)
      .def_readwrite("oy", &DenseSdfGridDevice::oy)
      .def_readwrite("oz", &DenseSdfGridDevice::oz)
      .def_property(
          "d_values",
          [](const DenseSdfGridDevice& g) { return (uintptr_t)g.d_values; },
          [](DenseSdfGridDevice& g, uintptr_t p) { g.d_values = (const float*)p; });
#endif
  py::class_<HipDualContouringBackend>(m, "HipDualContouringBackend")
      .def(py::init<int, int>(), py::arg("device") = 0, py::arg("precision") = 0)
      .def("extract",
           [](HipDualContouringBackend& self, const DenseSdfGrid& grid) {
             DualContouringStats stats;
             DualContouringMesh mesh = self.extract(grid, stats);
             return py::make_tuple(mesh, stats);
           })
      .def("extract_device",
           [](HipDualContouringBackend& self, const DenseSdfGridDevice& grid) {
             DualContouringStats stats;
             DualContouringMesh mesh = self.extract_device(grid, stats);
             return py::make_tuple(mesh, stats);
           })
      .def("postprocess", &HipDualContouringBackend::postprocess, py::arg("mesh"),
           py::arg("close_holes") = true, py::arg("remove_floaters") = true)
      .def("postprocess_last", &HipDualContouringBackend::postprocess_last,
           py::arg("close_holes") = true, py::arg("remove_floaters") = true)
      .def("last_vertices_f64", &HipDualContouringBackend::last_vertices_f64)
      .def("last_triangles", &HipDualContouringBackend::last_triangles)
      .def("set_option", &HipDualContouringBackend::set_option)
      .def_property_readonly("precision", &HipDualContouringBackend::precision);
  // bindings.cpp:229-262 — normal_mode as the integer value of the reference's
  // NormalComputationMode (0 FiniteDifference, 1 EdgeGradient, 2 PrecomputedGradient)
  py::class_<HipSparseDualContouringBackend>(m, "HipSparseDualContouringBackend")
      .def(py::init<int, int, int, bool, int>(), py::arg("brick_size") = 8,
           py::arg("normal_mode") = 0, py::arg("chunk_size") = 0,
           py::arg("multi_vertex_cells") = false, py::arg("device") = 0)
      .def_readwrite("brick_size", &HipSparseDualContouringBackend::brick_size)
      .def_readwrite("normal_mode", &HipSparseDualContouringBackend::normal_mode)
      .def_readwrite("chunk_size", &HipSparseDualContouringBackend::chunk_size)
      .def_readwrite("multi_vertex_cells", &HipSparseDualContouringBackend::multi_vertex_cells)
      .def("extract",
           [](HipSparseDualContouringBackend& self, const DenseSdfGrid& grid) {
             DualContouringStats stats;
             DualContouringMesh mesh = self.extract(grid, stats);
             return py::make_tuple(mesh, stats);
           })
      .def("extract_device",
           [](HipSparseDualContouringBackend& self, const DenseSdfGridDevice& grid) {
             DualContouringStats stats;
             DualContouringMesh mesh = self.extract_device(grid, stats);
             return py::make_tuple(mesh, stats);
           })
      .def("postprocess_last", &HipSparseDualContouringBackend::postprocess_last,
           py::arg("close_holes") = true, py::arg("remove_floaters") = true);
  py::class_<HipSparseMvdcDualContouringBackend, HipSparseDualContouringBackend>(
      m, "HipSparseMvdcDualContouringBackend")
      .def(py::init<int, int, int, int>(), py::arg("brick_size") = 8, py::arg("normal_mode") = 0,
           py::arg("chunk_size") = 0, py::arg("device") = 0);
  using Floats = py::array_t<float, py::array::c_style | py::array::forcecast>;
  using Ints = py::array_t<int, py::array::c_style | py::array::forcecast>;
  m.def("mesh_to_sdf_hip", [](Floats vertices, Ints faces, int nx, int ny, int nz, float ox,
                              float oy, float oz, float vx, float vy, float vz) {
    float ms = 0.0f;
    DenseSdfGrid grid = compute_mesh_sdf_hip(vertices.data(), (int)(vertices.size() / 3),
                                             faces.data(), (int)(faces.size() / 3), nx, ny, nz,
                                             ox, oy, oz, vx, vy, vz, &ms);
    return py::make_tuple(grid, ms);
  });
  m.def("mesh_to_sdf_device_hip", [](Floats vertices, Ints faces, int nx, int ny, int nz,
                                     float ox, float oy, float oz, float vx, float vy, float vz) {
    float ms = 0.0f;
    DenseSdfGridDevice grid = compute_mesh_sdf_device_hip(
        vertices.data(), (int)(vertices.size() / 3), faces.data(), (int)(faces.size() / 3), nx,
        ny, nz, ox, oy, oz, vx, vy, vz, &ms);
    return py::make_tuple(grid, ms);
  });
  m.def("free_mesh_sdf_device_hip", [](DenseSdfGridDevice& grid) { free_mesh_sdf_device_hip(grid); });
}
