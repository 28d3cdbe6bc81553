// dc_sparse.h — sparse brick dual contouring and its multi-vertex form on the GPU: the
// reference's CudaSparseDualContouringBackend and CudaSparseMvdcDualContouringBackend
// (src/cuda/dc_cuda_sparse.cu:1170-1485, declared include/dc_cuda.h:37-62).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "dc_hip.h"

class DcSparse {
 public:
  DcSparse() = default;
  ~DcSparse();
  DcSparse(const DcSparse&) = delete;
  DcSparse& operator=(const DcSparse&) = delete;

  // Extracts the grid (values: nx*ny*nz floats, x fastest, host or device memory as `on_device`
  // says) on stream s; the mesh stays in device memory (vertices(), faces()) until the next call.
  // Returns a dc_hip_status; on failure *err names the cause. Synchronises s.
  int run(const dc_hip_grid& grid, bool on_device, int brick_size, int normal_mode,
          bool multi_vertex, int cu_count, hipStream_t s, dc_hip_sparse_stats* st,
          std::string* err);

  bool have() const { return have_; }
  int64_t n_vertices() const { return nv_; }
  int64_t n_faces() const { return nf_; }
  const float* vertices() const { return static_cast<const float*>(b_[kVerts].p); }
  const int32_t* faces() const { return static_cast<const int32_t*>(b_[kFaces].p); }

 private:
  struct Buf {
    void* p = nullptr;
    size_t cap = 0;
  };
  enum {
    kUp,                   // the grid's device copy (host input)
    kPlanes,               // N, P and C sample bit planes
    kRows,                 // per brick row: the OR of N and of P over its sample rows
    kBricks,               // uint8 per brick: active
    kAct, kWpre,           // active-cell plane and its per-word ranks
    kScanSum, kScanBase,   // tiles of the active-plane scan
    kAlist,                // active cells in rank order
    kCtr,                  // counters
    kVerts,                // vertices (3 floats each)
    kSlotPos, kSlotNrm,    // multi-vertex form: two position and normal slots per active cell
    kVcnt, kVstart,        // multi-vertex form: vertices per active cell and their scan
    kFcnt, kFstart,        // quads per wave of active cells and their scan
    kStage,                // per-wave staging runs of quads
    kTsum,                 // tile sums of the uint32 scans
    kFaces,                // quads (4 int32 each)
    kNumBufs
  };
  int grow(int which, size_t bytes);
  int hip_error(hipError_t e, const char* what);

  Buf b_[kNumBufs];
  hipStream_t s_ = nullptr;
  int cu_ = 256;  // compute units of the device (grid-stride launch sizes)
  std::string* err_ = nullptr;
  uint32_t* h_ctr_ = nullptr;  // pinned host copy of the counters
  hipEvent_t ev_[20] = {};
  bool have_ = false;
  int64_t nv_ = 0, nf_ = 0;
};
