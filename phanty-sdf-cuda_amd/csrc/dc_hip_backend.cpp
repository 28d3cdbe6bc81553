// dc_hip_backend.cpp — IDualContouringBackend over the C ABI, plus DenseSdfGrid::sample.

#include "dc_hip_backend.h"

#include <algorithm>
#include <chrono>
#include <stdexcept>
#include <string>

// src/cpp/dc_backend.cpp:8-49: clamped trilinear sample.
float DenseSdfGrid::sample(float x, float y, float z) const {
  const float lx = (x - ox) / vx, ly = (y - oy) / vy, lz = (z - oz) / vz;
  const int x0 = std::max(0, std::min((int)lx, nx - 2));
  const int y0 = std::max(0, std::min((int)ly, ny - 2));
  const int z0 = std::max(0, std::min((int)lz, nz - 2));
  const int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
  const float tx = std::max(0.0f, std::min(1.0f, lx - x0));
  const float ty = std::max(0.0f, std::min(1.0f, ly - y0));
  const float tz = std::max(0.0f, std::min(1.0f, lz - z0));
  auto at = [&](int i, int j, int k) { return values[i + nx * (j + ny * k)]; };
  const float c00 = at(x0, y0, z0) * (1.0f - tx) + at(x1, y0, z0) * tx;
  const float c01 = at(x0, y0, z1) * (1.0f - tx) + at(x1, y0, z1) * tx;
  const float c10 = at(x0, y1, z0) * (1.0f - tx) + at(x1, y1, z0) * tx;
  const float c11 = at(x0, y1, z1) * (1.0f - tx) + at(x1, y1, z1) * tx;
  const float c0 = c00 * (1.0f - ty) + c10 * ty;
  const float c1 = c01 * (1.0f - ty) + c11 * ty;
  return c0 * (1.0f - tz) + c1 * tz;
}

HipDualContouringBackend::HipDualContouringBackend(int device, int precision)
    : precision_(precision) {
  const int rc = dc_hip_create(device, &ctx_);
  if (rc != DC_HIP_OK)
    throw std::runtime_error(std::string("HipDualContouringBackend: ") +
                             dc_hip_status_string(rc));
}

HipDualContouringBackend::~HipDualContouringBackend() { dc_hip_destroy(ctx_); }

void HipDualContouringBackend::check(int rc) {
  if (rc != DC_HIP_OK)
    throw std::runtime_error(std::string("HipDualContouringBackend: ") +
                             dc_hip_status_string(rc) + ": " + dc_hip_last_error(ctx_));
}

void HipDualContouringBackend::set_option(const std::string& key, int64_t value) {
  check(dc_hip_set_option(ctx_, key.c_str(), value));
}

DualContouringMesh HipDualContouringBackend::download(const dc_hip_stats& st,
                                                      DualContouringStats& stats) {
  const auto t0 = std::chrono::steady_clock::now();
  DualContouringMesh mesh;
  mesh.vertices.resize(3 * (size_t)st.vertex_count);
  mesh.faces.resize(4 * (size_t)st.face_count);
  check(dc_hip_copy_result(ctx_, mesh.vertices.empty() ? nullptr : mesh.vertices.data(), nullptr,
                           mesh.faces.empty() ? nullptr : mesh.faces.data()));
  const float dl =
      std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  stats = DualContouringStats();
  stats.backend = "hip";
  stats.nx = st.nx;
  stats.ny = st.ny;
  stats.nz = st.nz;
  stats.total_cells = (int)st.total_cells;
  stats.active_cells = (int)st.active_cells;
  stats.upload_ms = st.upload_ms;
  stats.marking_ms = st.marking_ms;
  stats.active_cell_marking_ms = st.marking_ms;
  stats.compaction_ms = st.compaction_ms;
  stats.active_cell_compaction_ms = st.compaction_ms;
  stats.qef_ms = st.qef_ms;
  stats.face_prefix_sum_ms = st.face_prefix_sum_ms;
  stats.face_fill_ms = st.face_fill_ms;
  stats.face_emission_ms = st.face_emission_ms;
  stats.download_ms = dl;
  stats.total_ms = st.total_ms + dl;
  stats.vertex_count = (int)st.vertex_count;
  stats.face_count = (int)st.face_count;
  stats.qef_fallback_count = (int)st.qef_fallback_count;
  stats.invalid_count = (int)st.invalid_count;
  return mesh;
}

DualContouringMesh HipDualContouringBackend::extract(const DenseSdfGrid& grid,
                                                     DualContouringStats& stats) {
  if ((size_t)grid.nx * grid.ny * grid.nz != grid.values.size())
    throw std::runtime_error("HipDualContouringBackend: values size does not match nx*ny*nz");
  dc_hip_grid g{grid.nx, grid.ny, grid.nz, grid.vx, grid.vy, grid.vz,
                grid.ox, grid.oy, grid.oz, grid.values.data()};
  dc_hip_stats st{};
  check(dc_hip_extract(ctx_, &g, precision_, &st));
  return download(st, stats);
}

DualContouringMesh HipDualContouringBackend::extract_device(const DenseSdfGridDevice& grid,
                                                            DualContouringStats& stats) {
  dc_hip_grid g{grid.nx, grid.ny, grid.nz, grid.vx, grid.vy, grid.vz,
                grid.ox, grid.oy, grid.oz, grid.d_values};
  dc_hip_stats st{};
  check(dc_hip_extract_device(ctx_, &g, precision_, &st));
  return download(st, stats);
}

std::vector<double> HipDualContouringBackend::last_vertices_f64() {
  int64_t nv = 0, nf = 0;
  check(dc_hip_result_counts(ctx_, &nv, &nf));
  std::vector<double> v(3 * (size_t)nv);
  if (nv) check(dc_hip_copy_result(ctx_, nullptr, v.data(), nullptr));
  return v;
}

std::vector<int> HipDualContouringBackend::last_triangles() {
  int64_t nv = 0, nf = 0;
  check(dc_hip_result_counts(ctx_, &nv, &nf));
  std::vector<int> t(6 * (size_t)nf);
  if (nf) check(dc_hip_triangulate(ctx_, t.data()));
  return t;
}
