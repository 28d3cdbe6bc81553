// dc_post.h — mesh post-processing on the GPU: vertex weld, floater removal and hole closing, the
// steps of the reference's postprocess_mesh (src/cpp/dc_backend.cpp:127-558).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "dc_hip.h"

class DcPost {
 public:
  DcPost() = default;
  ~DcPost();
  DcPost(const DcPost&) = delete;
  DcPost& operator=(const DcPost&) = delete;

  // Post-processes the quad mesh (vertices: 3 * nv floats, faces: 4 * nf int32; host or device
  // memory) on stream s. The result stays in this object (vertices(), faces(), counts). Returns a
  // dc_hip_status; on failure *err names the cause. Synchronises s.
  int run(const float* vertices, int64_t nv, const int32_t* faces, int64_t nf, bool close_holes,
          bool remove_floaters, int cu_count, hipStream_t s, dc_hip_post_stats* st,
          std::string* err);

  int64_t n_vertices() const { return nv_out_; }
  int64_t n_faces() const { return nf_out_; }
  const float* vertices() const { return out_v_; }
  const int32_t* faces() const { return out_f_; }
  bool have() const { return have_; }

 private:
  struct Buf {
    void* p = nullptr;
    size_t cap = 0;
  };
  enum {
    kInV, kInF,            // the input mesh (device copy)
    kWV, kWF,              // first weld's result
    kW2V, kW2F,            // second weld's result
    kPart, kTable, kSlot, kRemap, kFTmp,  // weld scratch
    kFlag, kIdx, kTSum, kCnt,             // flags, their exclusive scan, tile sums, totals
    kETab, kECnt, kEPair, kESlot,         // undirected edge table
    kParent, kCSize, kBest, kKeep,        // components
    kBnd, kAF, kNewV, kFirst, kRV, kRF,   // boundary edges, faces before renumbering, results
    kNumBufs
  };
  int grow(int which, size_t bytes);
  int hip_error(hipError_t e, const char* what);
  // enqueues an exclusive scan of n uint32 flags into kIdx (total into kCnt[slot]).
  int scan(const uint32_t* flags, int64_t n, uint32_t* out, int slot);
  int read_counts(uint32_t* h, int n);
  // weld of (v, nv, f, nf) into buffers ov / of; returns the counts.
  int weld(const float* v, int64_t nv, const int32_t* f, int64_t nf, int ov, int of,
           int64_t* nvo, int64_t* nfo);

  Buf b_[kNumBufs];
  hipStream_t s_ = nullptr;
  int cu_ = 256;
  std::string* err_ = nullptr;
  bool have_ = false;
  int64_t nv_out_ = 0, nf_out_ = 0;
  float* out_v_ = nullptr;
  int32_t* out_f_ = nullptr;
};
