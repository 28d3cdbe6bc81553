// dc_flagscan.h — exclusive scan of uint32 flags or counts in three launches: tile sums, a
// one-block scan of the tile sums (writing the total), then each tile scanned again with its
// base. Used by the mesh post-processing and the mesh-to-SDF passes (not the extraction, whose
// scans work on bit words).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dc_common.h"

// The kernels have internal linkage: every translation unit that includes this header gets its
// own copies.
namespace dcscan {
namespace {

using dccommon::wave_scan_incl;

constexpr int kThreads = 256;
constexpr int kPer = 8;                     // values per thread of a scan tile
constexpr int64_t kTile = kThreads * kPer;  // values per scan tile

__device__ __forceinline__ uint32_t block_excl(uint32_t x, uint32_t* s_w, uint32_t* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t incl = wave_scan_incl(x, lane);
  if (lane == 63) s_w[w] = incl;
  __syncthreads();
  uint32_t base = 0, tot = 0;
#pragma unroll
  for (int q = 0; q < kThreads / 64; ++q) {
    if (q < w) base += s_w[q];
    tot += s_w[q];
  }
  __syncthreads();
  *total = tot;
  return base + incl - x;
}

__global__ void __launch_bounds__(kThreads)
    k_scan_sum(const uint32_t* __restrict__ in, int64_t n, uint32_t* __restrict__ tsum) {
  __shared__ uint32_t s_w[kThreads / 64];
  const int64_t base = (int64_t)blockIdx.x * kTile + threadIdx.x * kPer;
  uint32_t x = 0;
#pragma unroll
  for (int k = 0; k < kPer; ++k)
    if (base + k < n) x += in[base + k];
  uint32_t tot;
  block_excl(x, s_w, &tot);
  if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kThreads)
    k_scan_tiles(uint32_t* __restrict__ tsum, int64_t nt, uint32_t* __restrict__ total) {
  __shared__ uint32_t s_w[kThreads / 64];
  uint32_t carry = 0;
  for (int64_t off = 0; off < nt; off += kThreads) {
    const int64_t q = off + threadIdx.x;
    const uint32_t x = q < nt ? tsum[q] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl(x, s_w, &tot);
    if (q < nt) tsum[q] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(kThreads)
    k_scan_apply(const uint32_t* __restrict__ in, int64_t n, const uint32_t* __restrict__ tbase,
                 uint32_t* __restrict__ out) {
  __shared__ uint32_t s_w[kThreads / 64];
  const int64_t base = (int64_t)blockIdx.x * kTile + threadIdx.x * kPer;
  uint32_t v[kPer], x = 0;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    v[k] = base + k < n ? in[base + k] : 0u;
    x += v[k];
  }
  uint32_t tot;
  uint32_t run = tbase[blockIdx.x] + block_excl(x, s_w, &tot);
#pragma unroll
  for (int k = 0; k < kPer; ++k)
    if (base + k < n) {
      out[base + k] = run;
      run += v[k];
    }
}


inline int64_t scan_tiles(int64_t n) { return n > 0 ? (n + kTile - 1) / kTile : 0; }

// out[i] = in[0] + ... + in[i-1] (n values); *total = their sum. tsum: scan_tiles(n) entries
// (at least one).
inline hipError_t exclusive_scan(const uint32_t* in, int64_t n, uint32_t* out, uint32_t* tsum,
                                 uint32_t* total, hipStream_t s) {
  const int64_t nt = scan_tiles(n);
  if (nt > 0) hipLaunchKernelGGL(k_scan_sum, dim3((unsigned)nt), dim3(kThreads), 0, s, in, n, tsum);
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kThreads), 0, s, tsum, nt, total);
  if (nt > 0)
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nt), dim3(kThreads), 0, s, in, n, tsum, out);
  return hipGetLastError();
}

}  // namespace
}  // namespace dcscan
