// dc_m2s.h — mesh to signed distance grid on the GPU: the reference's mesh-to-SDF producer,
// compute_mesh_sdf_device_cuda (src/cuda/mesh_to_sdf_cuda.cu:740-973), the step before the
// extraction for mesh inputs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "dc_hip.h"

class DcM2s {
 public:
  DcM2s() = default;
  ~DcM2s();
  DcM2s(const DcM2s&) = delete;
  DcM2s& operator=(const DcM2s&) = delete;

  // Samples the signed distance of the triangle mesh (vertices: 3 * nv floats, triangles:
  // 3 * nt int32; host or device memory) on the grid `dims` / `origin` / `spacing` into `out`
  // (dims[0] * dims[1] * dims[2] floats, x fastest; host or device memory) on stream s. Returns a
  // dc_hip_status; on failure *err names the cause. Synchronises s.
  int run(const float* vertices, int64_t nv, const int32_t* triangles, int64_t nt,
          const int32_t dims[3], const float origin[3], const float spacing[3],
          bool voxelize_first, int32_t voxel_res, float* out, int cu_count, hipStream_t s,
          dc_hip_m2s_stats* st, std::string* err);

 private:
  struct Buf {
    void* p = nullptr;
    size_t cap = 0;
  };
  enum {
    kInV, kInT,           // the input mesh (device copy)
    kTri, kBox, kCnt,     // per triangle: record, bin box, bin count
    kOff, kTSum, kCtr,    // pair offsets, scan tile sums, counters
    kKeyA, kValA, kKeyB, kValB, kHist,  // triangle-bin pairs and the radix sort's buffers
    kBinCnt, kBinStart, kBinAct, kActIdx, kActList,  // per bin
    kAxis,                // per-axis sample-to-bin and bin-to-sample tables
    kP, kE,               // flood masks (samples or voxels)
    kShell,               // voxelize: shell bits
    kVals, kTmp,          // the grid (when `out` is host memory) and the relaxation's second grid
    kNumBufs
  };
  int grow(int which, size_t bytes);
  int hip_error(hipError_t e, const char* what);
  int flood(uint64_t* P, uint64_t* E, int nx, int ny, int nz, int* passes);

  Buf b_[kNumBufs];
  hipStream_t s_ = nullptr;
  int cu_ = 256;
  std::string* err_ = nullptr;
  std::vector<int32_t> axis_host_;  // per-axis tables, kept until the next call (async upload)
};
