// dc_backend.h (standalone) — grid, mesh, stats and backend interface of the dense
// dual-contouring path, restated API-compatibly from the reference (include/dc_backend.h:1-89).
//
// Only this repository's own builds (its dc_cli, the Python-free C++ users) put this directory on
// the include path. A build inside the reference tree includes the reference's own
// include/dc_backend.h instead: dc_hip_backend.h names <dc_backend.h> and the include path picks
// one, so the two never meet in one translation unit. Only the types on the dense path are
// declared; the CPU backend stays in the reference.
#pragma once

#include <memory>
#include <string>
#include <vector>

class ISdfSampler {
 public:
  virtual ~ISdfSampler() = default;
  virtual float sample(float x, float y, float z) const = 0;
};

// include/dc_backend.h:13-21: values[i + nx*(j + ny*k)], sample (i,j,k) at (ox + i*vx, ...).
class DenseSdfGrid : public ISdfSampler {
 public:
  int nx = 0, ny = 0, nz = 0;
  float vx = 0.f, vy = 0.f, vz = 0.f;
  float ox = 0.f, oy = 0.f, oz = 0.f;
  std::vector<float> values;

  // Trilinear sample with the clamping of src/cpp/dc_backend.cpp:8-49.
  float sample(float x, float y, float z) const override;
};

// include/dc_backend.h:23-26
struct DualContouringMesh {
  std::vector<float> vertices;  // 3 per vertex
  std::vector<int> faces;       // 4 per quad, 0-based
};

// include/dc_backend.h:44-78 (field for field)
struct DualContouringStats {
  std::string backend;
  int nx = 0, ny = 0, nz = 0;
  int total_cells = 0;
  int active_cells = 0;
  int total_bricks = 0;
  int active_bricks = 0;
  float upload_ms = 0.f;
  float marking_ms = 0.f;
  float compaction_ms = 0.f;
  float active_brick_marking_ms = 0.f;
  float active_brick_compaction_ms = 0.f;
  float active_cell_marking_ms = 0.f;
  float active_cell_compaction_ms = 0.f;
  float qef_ms = 0.f;
  float face_emission_ms = 0.f;
  float face_count_ms = 0.f;
  float face_prefix_sum_ms = 0.f;
  float face_fill_ms = 0.f;
  float download_ms = 0.f;
  float total_ms = 0.f;
  int vertex_count = 0;
  int face_count = 0;
  int qef_fallback_count = 0;
  int clamp_count = 0;
  int invalid_count = 0;
  int ambiguous_cells = 0;
  int multi_vertex_cells = 0;
  int one_vertex_cells = 0;
  int two_vertex_cells = 0;
  int split_rejection_count = 0;
  int bad_qef_count = 0;
  int faces_skipped_due_to_missing_cluster = 0;
};

// include/dc_backend.h:80-84
class IDualContouringBackend {
 public:
  virtual ~IDualContouringBackend() = default;
  virtual DualContouringMesh extract(const DenseSdfGrid& grid, DualContouringStats& stats) = 0;
};
