// dc_hip_backend.h — C++ backend class over the C ABI (dc_hip.h); takes the place of
// CudaDualContouringBackend (reference include/dc_cuda.h:31-35) behind IDualContouringBackend.
//
// The grid, mesh and stats types come from <dc_backend.h>: inside the reference tree that is the
// reference's include/dc_backend.h (put this directory after the reference's include directory);
// this repository's own builds add include/standalone, which restates it. DenseSdfGridDevice is
// the reference's (include/dc_cuda.h:24-29) whenever the build enables the CUDA backends
// (DC_ENABLE_CUDA, CMakeLists.txt:52-72), and a field-for-field equal struct otherwise.
// This header defines nothing the reference's dc_core also defines (DenseSdfGrid::sample lives in
// the reference's src/cpp/dc_backend.cpp:8-49, or in this repository's dc_standalone.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include <dc_backend.h>

#include "dc_hip.h"

#ifdef DC_ENABLE_CUDA
#include <dc_cuda.h>
#else
// include/dc_cuda.h:24-29: a grid whose values already live in device memory (caller-owned).
struct DenseSdfGridDevice {
  int nx, ny, nz;
  float vx, vy, vz;
  float ox, oy, oz;
  const float* d_values;
};
#endif

class HipDualContouringBackend : public IDualContouringBackend {
 public:
  // Throws std::runtime_error when no HIP device is available.
  explicit HipDualContouringBackend(int device = 0, int precision = DC_HIP_PLAIN_FP64);
  ~HipDualContouringBackend() override;
  HipDualContouringBackend(const HipDualContouringBackend&) = delete;
  HipDualContouringBackend& operator=(const HipDualContouringBackend&) = delete;

  // dc_cuda.cu:614-649 — host grid in, host mesh out.
  DualContouringMesh extract(const DenseSdfGrid& grid, DualContouringStats& stats) override;
  // dc_cuda.cu:651-847 — device grid in, host mesh out.
  DualContouringMesh extract_device(const DenseSdfGridDevice& grid, DualContouringStats& stats);
  // FP64 vertices of the last extraction (plain FP64 mode).
  std::vector<double> last_vertices_f64();
  // The last extraction's quads as triangles (3 ints each), split and oriented like
  // triangulate_v1 (src/cpp/contouring.cpp:88-113).
  std::vector<int> last_triangles();
  // postprocess_mesh (src/cpp/dc_backend.cpp:217-558, declared include/dc_backend.h:28) on the
  // GPU: weld, largest component (remove_floaters), closed boundary loops (close_holes); the
  // result equals the reference's.
  DualContouringMesh postprocess(const DualContouringMesh& mesh, bool close_holes,
                                 bool remove_floaters);
  // The same on the last extraction's mesh, without copying it to the host first (single grid).
  DualContouringMesh postprocess_last(bool close_holes, bool remove_floaters);
  // compute_mesh_sdf_cuda / compute_mesh_sdf_device_cuda (include/mesh_to_sdf.h:6-26) on this
  // backend's GPU: the triangle mesh's signed distance sampled on the grid, equal to the
  // reference's result (dc_hip_mesh_to_sdf). The device form writes into `values`, a device
  // buffer of nx*ny*nz floats the caller owns.
  DenseSdfGrid mesh_to_sdf(const float* vertices, int num_vertices, const int* faces,
                           int num_faces, int nx, int ny, int nz, float ox, float oy, float oz,
                           float vx, float vy, float vz, float* out_ms = nullptr,
                           bool voxelize_first = false, int voxel_res = 256);
  void mesh_to_sdf_device(const float* vertices, int num_vertices, const int* faces,
                          int num_faces, int nx, int ny, int nz, float ox, float oy, float oz,
                          float vx, float vy, float vz, float* values, float* out_ms = nullptr,
                          bool voxelize_first = false, int voxel_res = 256);
  void set_option(const std::string& key, int64_t value);
  int precision() const { return precision_; }
  dc_hip_context* context() { return ctx_; }

 private:
  DualContouringMesh post_copy(int64_t nv, int64_t nf);
  DualContouringMesh download(const dc_hip_stats& st, DualContouringStats& stats);
  void check(int rc);
  dc_hip_context* ctx_ = nullptr;
  int precision_;
};

// include/dc_cuda.h:37-56 (CudaSparseDualContouringBackend, src/cuda/dc_cuda_sparse.cu:1134-1485):
// the sparse brick extraction (dc_hip_extract_sparse) behind IDualContouringBackend, with the
// reference's public fields. normal_mode takes the values of the reference's NormalComputationMode
// (dc_hip_normal_mode). The mesh and every counter equal the reference's; chunk_size has no
// effect, as in the reference. Throws std::runtime_error on failure.
class HipSparseDualContouringBackend : public IDualContouringBackend {
 public:
  explicit HipSparseDualContouringBackend(int brick_size = 8,
                                          int normal_mode = DC_HIP_NORMAL_FINITE_DIFFERENCE,
                                          int chunk_size = 0, bool multi_vertex_cells = false,
                                          int device = 0);
  ~HipSparseDualContouringBackend() override;
  HipSparseDualContouringBackend(const HipSparseDualContouringBackend&) = delete;
  HipSparseDualContouringBackend& operator=(const HipSparseDualContouringBackend&) = delete;

  // dc_cuda_sparse.cu:1134-1168 — host grid in, host mesh out.
  DualContouringMesh extract(const DenseSdfGrid& grid, DualContouringStats& stats) override;
  // dc_cuda_sparse.cu:1170-1485 — device grid in, host mesh out.
  DualContouringMesh extract_device(const DenseSdfGridDevice& grid, DualContouringStats& stats);
  // postprocess_mesh (src/cpp/dc_backend.cpp:217-558) on the last sparse mesh, on the GPU,
  // without copying the mesh to the host first.
  DualContouringMesh postprocess_last(bool close_holes, bool remove_floaters);
  dc_hip_context* context() { return ctx_; }

  int brick_size;
  int normal_mode;
  int chunk_size;
  bool multi_vertex_cells;

 private:
  DualContouringMesh finish(const dc_hip_sparse_stats& st, DualContouringStats& stats);
  void check(int rc);
  dc_hip_context* ctx_ = nullptr;
};

// include/dc_cuda.h:58-62 (CudaSparseMvdcDualContouringBackend): the multi-vertex form.
class HipSparseMvdcDualContouringBackend : public HipSparseDualContouringBackend {
 public:
  explicit HipSparseMvdcDualContouringBackend(int brick_size = 8,
                                              int normal_mode = DC_HIP_NORMAL_FINITE_DIFFERENCE,
                                              int chunk_size = 0, int device = 0)
      : HipSparseDualContouringBackend(brick_size, normal_mode, chunk_size, true, device) {}
};

// include/mesh_to_sdf.h:6-28 with the backend's name: the mesh's signed distance grid on the host
// (compute_mesh_sdf_hip), or on the device (compute_mesh_sdf_device_hip: d_values allocated with
// dc_hip_malloc, released by free_mesh_sdf_device_hip). out_ms receives the call's wall time.
// They run on device 0 through one process-wide context (calls are serialised). Throw
// std::runtime_error on failure.
DenseSdfGrid compute_mesh_sdf_hip(const float* vertices, int num_vertices, const int* faces,
                                  int num_faces, int nx, int ny, int nz, float ox, float oy,
                                  float oz, float vx, float vy, float vz, float* out_ms = nullptr,
                                  bool voxelize_first = false, int voxel_res = 256);
DenseSdfGridDevice compute_mesh_sdf_device_hip(const float* vertices, int num_vertices,
                                               const int* faces, int num_faces, int nx, int ny,
                                               int nz, float ox, float oy, float oz, float vx,
                                               float vy, float vz, float* out_ms = nullptr,
                                               bool voxelize_first = false, int voxel_res = 256);
void free_mesh_sdf_device_hip(DenseSdfGridDevice& device_grid);
