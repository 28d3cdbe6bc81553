// dc_hip_backend.cpp — IDualContouringBackend over the C ABI. Compiles against the reference's
// include/dc_backend.h (inside the reference tree) or this repository's standalone restatement,
// and defines only the members of HipDualContouringBackend and HipSparseDualContouringBackend and
// the mesh-to-SDF free functions.

#include "dc_hip_backend.h"

#include <algorithm>
#include <chrono>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>

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

DualContouringMesh HipDualContouringBackend::post_copy(int64_t nv, int64_t nf) {
  DualContouringMesh m;
  m.vertices.resize(3 * (size_t)nv);
  m.faces.resize(4 * (size_t)nf);
  check(dc_hip_postprocess_copy(ctx_, m.vertices.data(), m.faces.data()));
  return m;
}

DualContouringMesh HipDualContouringBackend::postprocess(const DualContouringMesh& mesh,
                                                         bool close_holes, bool remove_floaters) {
  int64_t nv = 0, nf = 0;
  check(dc_hip_postprocess(ctx_, mesh.vertices.data(), (int64_t)(mesh.vertices.size() / 3),
                           mesh.faces.data(), (int64_t)(mesh.faces.size() / 4), close_holes ? 1 : 0,
                           remove_floaters ? 1 : 0, &nv, &nf, nullptr));
  return post_copy(nv, nf);
}

DualContouringMesh HipDualContouringBackend::postprocess_last(bool close_holes,
                                                              bool remove_floaters) {
  int64_t nv = 0, nf = 0;
  check(dc_hip_postprocess(ctx_, nullptr, 0, nullptr, 0, close_holes ? 1 : 0,
                           remove_floaters ? 1 : 0, &nv, &nf, nullptr));
  return post_copy(nv, nf);
}

void HipDualContouringBackend::mesh_to_sdf_device(const float* vertices, int num_vertices,
                                                  const int* faces, int num_faces, int nx, int ny,
                                                  int nz, float ox, float oy, float oz, float vx,
                                                  float vy, float vz, float* values, float* out_ms,
                                                  bool voxelize_first, int voxel_res) {
  dc_hip_m2s_stats st{};
  check(dc_hip_mesh_to_sdf(ctx_, vertices, num_vertices, faces, num_faces, nx, ny, nz, ox, oy, oz,
                           vx, vy, vz, voxelize_first ? 1 : 0, voxel_res, values, &st));
  if (out_ms) *out_ms = st.total_ms;
}

DenseSdfGrid HipDualContouringBackend::mesh_to_sdf(const float* vertices, int num_vertices,
                                                   const int* faces, int num_faces, int nx, int ny,
                                                   int nz, float ox, float oy, float oz, float vx,
                                                   float vy, float vz, float* out_ms,
                                                   bool voxelize_first, int voxel_res) {
  DenseSdfGrid g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.ox = ox;
  g.oy = oy;
  g.oz = oz;
  g.vx = vx;
  g.vy = vy;
  g.vz = vz;
  if (nx < 1 || ny < 1 || nz < 1)
    throw std::runtime_error("HipDualContouringBackend: nx, ny and nz must be >= 1");
  g.values.resize((size_t)nx * ny * nz);
  mesh_to_sdf_device(vertices, num_vertices, faces, num_faces, nx, ny, nz, ox, oy, oz, vx, vy, vz,
                     g.values.data(), out_ms, voxelize_first, voxel_res);
  return g;
}

namespace {
std::mutex g_m2s_mutex;
std::unique_ptr<HipDualContouringBackend> g_m2s_backend;

HipDualContouringBackend& m2s_backend() {
  if (!g_m2s_backend) g_m2s_backend.reset(new HipDualContouringBackend(0));
  return *g_m2s_backend;
}
}  // namespace

DenseSdfGrid compute_mesh_sdf_hip(const float* vertices, int num_vertices, const int* faces,
                                  int num_faces, int nx, int ny, int nz, float ox, float oy,
                                  float oz, float vx, float vy, float vz, float* out_ms,
                                  bool voxelize_first, int voxel_res) {
  std::lock_guard<std::mutex> lock(g_m2s_mutex);
  return m2s_backend().mesh_to_sdf(vertices, num_vertices, faces, num_faces, nx, ny, nz, ox, oy,
                                   oz, vx, vy, vz, out_ms, voxelize_first, voxel_res);
}

DenseSdfGridDevice compute_mesh_sdf_device_hip(const float* vertices, int num_vertices,
                                               const int* faces, int num_faces, int nx, int ny,
                                               int nz, float ox, float oy, float oz, float vx,
                                               float vy, float vz, float* out_ms,
                                               bool voxelize_first, int voxel_res) {
  std::lock_guard<std::mutex> lock(g_m2s_mutex);
  HipDualContouringBackend& be = m2s_backend();
  if (nx < 1 || ny < 1 || nz < 1)
    throw std::runtime_error("compute_mesh_sdf_device_hip: nx, ny and nz must be >= 1");
  void* p = nullptr;
  const int rc = dc_hip_malloc(be.context(), (int64_t)nx * ny * nz * 4, &p);
  if (rc != DC_HIP_OK)
    throw std::runtime_error(std::string("compute_mesh_sdf_device_hip: ") +
                             dc_hip_status_string(rc) + ": " + dc_hip_last_error(be.context()));
  try {
    be.mesh_to_sdf_device(vertices, num_vertices, faces, num_faces, nx, ny, nz, ox, oy, oz, vx,
                          vy, vz, static_cast<float*>(p), out_ms, voxelize_first, voxel_res);
  } catch (...) {
    dc_hip_free(p);
    throw;
  }
  DenseSdfGridDevice g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.vx = vx;
  g.vy = vy;
  g.vz = vz;
  g.ox = ox;
  g.oy = oy;
  g.oz = oz;
  g.d_values = static_cast<const float*>(p);
  return g;
}

void free_mesh_sdf_device_hip(DenseSdfGridDevice& device_grid) {
  if (device_grid.d_values) {
    dc_hip_free(const_cast<float*>(device_grid.d_values));
    device_grid.d_values = nullptr;
  }
}

HipSparseDualContouringBackend::HipSparseDualContouringBackend(int brick_size_, int normal_mode_,
                                                               int chunk_size_,
                                                               bool multi_vertex_cells_,
                                                               int device)
    : brick_size(brick_size_),
      normal_mode(normal_mode_),
      chunk_size(chunk_size_),
      multi_vertex_cells(multi_vertex_cells_) {
  const int rc = dc_hip_create(device, &ctx_);
  if (rc != DC_HIP_OK)
    throw std::runtime_error(std::string("HipSparseDualContouringBackend: ") +
                             dc_hip_status_string(rc));
}

HipSparseDualContouringBackend::~HipSparseDualContouringBackend() { dc_hip_destroy(ctx_); }

void HipSparseDualContouringBackend::check(int rc) {
  if (rc != DC_HIP_OK)
    throw std::runtime_error(std::string("HipSparseDualContouringBackend: ") +
                             dc_hip_status_string(rc) + ": " + dc_hip_last_error(ctx_));
}

DualContouringMesh HipSparseDualContouringBackend::finish(const dc_hip_sparse_stats& st,
                                                          DualContouringStats& stats) {
  const auto t0 = std::chrono::steady_clock::now();
  DualContouringMesh mesh;
  mesh.vertices.resize(3 * (size_t)st.vertex_count);
  mesh.faces.resize(4 * (size_t)st.face_count);
  check(dc_hip_sparse_copy_result(ctx_, mesh.vertices.empty() ? nullptr : mesh.vertices.data(),
                                  mesh.faces.empty() ? nullptr : mesh.faces.data()));
  const float dl =
      std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  // dc_cuda_sparse.cu:1436-1482
  stats = DualContouringStats();
  stats.backend = multi_vertex_cells ? "hip-sparse-mvdc" : "hip-sparse";
  stats.nx = st.nx;
  stats.ny = st.ny;
  stats.nz = st.nz;
  stats.total_cells = (int)st.total_cells;
  stats.active_cells = (int)st.active_cells;
  stats.total_bricks = (int)st.total_bricks;
  stats.active_bricks = (int)st.active_bricks;
  stats.upload_ms = st.upload_ms;
  stats.active_brick_marking_ms = st.active_brick_marking_ms;
  stats.active_brick_compaction_ms = st.active_brick_compaction_ms;
  stats.active_cell_marking_ms = st.active_cell_marking_ms;
  stats.active_cell_compaction_ms = st.active_cell_compaction_ms;
  stats.marking_ms = st.marking_ms;
  stats.compaction_ms = st.compaction_ms;
  stats.qef_ms = st.qef_ms;
  stats.face_count_ms = st.face_count_ms;
  stats.face_prefix_sum_ms = st.face_prefix_sum_ms;
  stats.face_fill_ms = st.face_fill_ms;
  stats.face_emission_ms = st.face_emission_ms;
  stats.download_ms = dl;
  stats.total_ms = st.total_ms + dl;
  stats.vertex_count = (int)st.vertex_count;
  stats.face_count = (int)st.face_count;
  stats.qef_fallback_count = (int)st.qef_fallback_count;
  stats.clamp_count = (int)st.clamp_count;
  stats.invalid_count = (int)st.invalid_count;
  stats.ambiguous_cells = (int)st.ambiguous_cells;
  stats.multi_vertex_cells = (int)st.multi_vertex_cells;
  stats.one_vertex_cells = (int)st.one_vertex_cells;
  stats.two_vertex_cells = (int)st.two_vertex_cells;
  stats.split_rejection_count = (int)st.split_rejection_count;
  stats.bad_qef_count = (int)st.bad_qef_count;
  stats.faces_skipped_due_to_missing_cluster = (int)st.faces_skipped_due_to_missing_cluster;
  return mesh;
}

DualContouringMesh HipSparseDualContouringBackend::extract(const DenseSdfGrid& grid,
                                                           DualContouringStats& stats) {
  if ((size_t)grid.nx * grid.ny * grid.nz != grid.values.size())
    throw std::runtime_error("HipSparseDualContouringBackend: values size does not match nx*ny*nz");
  dc_hip_grid g{grid.nx, grid.ny, grid.nz, grid.vx, grid.vy, grid.vz,
                grid.ox, grid.oy, grid.oz, grid.values.data()};
  dc_hip_sparse_stats st{};
  check(dc_hip_extract_sparse(ctx_, &g, brick_size, normal_mode, multi_vertex_cells ? 1 : 0, &st));
  return finish(st, stats);
}

DualContouringMesh HipSparseDualContouringBackend::extract_device(const DenseSdfGridDevice& grid,
                                                                  DualContouringStats& stats) {
  dc_hip_grid g{grid.nx, grid.ny, grid.nz, grid.vx, grid.vy, grid.vz,
                grid.ox, grid.oy, grid.oz, grid.d_values};
  dc_hip_sparse_stats st{};
  check(dc_hip_extract_sparse_device(ctx_, &g, brick_size, normal_mode,
                                     multi_vertex_cells ? 1 : 0, &st));
  return finish(st, stats);
}

DualContouringMesh HipSparseDualContouringBackend::postprocess_last(bool close_holes,
                                                                    bool remove_floaters) {
  int64_t nv = 0, nf = 0;
  const float* dv = nullptr;
  const int32_t* df = nullptr;
  check(dc_hip_sparse_result_device(ctx_, &nv, &nf, &dv, &df));
  if (nv == 0 && nf == 0) return DualContouringMesh();  // an empty mesh stays empty
  int64_t onv = 0, onf = 0;
  check(dc_hip_postprocess(ctx_, dv, nv, df, nf, close_holes ? 1 : 0, remove_floaters ? 1 : 0,
                           &onv, &onf, nullptr));
  DualContouringMesh m;
  m.vertices.resize(3 * (size_t)onv);
  m.faces.resize(4 * (size_t)onf);
  check(dc_hip_postprocess_copy(ctx_, m.vertices.empty() ? nullptr : m.vertices.data(),
                                m.faces.empty() ? nullptr : m.faces.data()));
  return m;
}
