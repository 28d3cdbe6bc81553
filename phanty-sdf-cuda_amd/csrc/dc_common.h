// dc_common.h — small device helpers shared by the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dc_internal.h"

// Pointers in the global address space. The sample pointer reaches the kernels inside the grid
// descriptor (loaded from memory), where the compiler cannot tell that it is global and would
// issue flat loads (slower, and counted against the LDS/scalar wait counter as well).
typedef const __attribute__((address_space(1))) float* gfloat;

namespace dccommon {

__device__ __forceinline__ gfloat to_global(const float* p) { return (gfloat)p; }

// Largest b in [0, n) with bounds[b] <= x (bounds ascending, bounds[0] = 0): the grid of a batch
// that owns item x.
__device__ __forceinline__ int find_bucket(const int64_t* __restrict__ bounds, int n, int64_t x) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (bounds[mid] <= x)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Workgroup b runs on XCD b % 8. Remap so each XCD gets a contiguous range of tiles, keeping
// neighbouring tiles (which share halo rows) in one L2.
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t nb) {
  const int64_t q = nb >> 3, r = nb & 7, x = b & 7, s = b >> 3;
  return x * q + (x < r ? x : r) + s;
}

// This block's share [lo, hi) of n items split into gridDim.x contiguous shares, the shares handed
// out XCD-aware (the blocks of one XCD take consecutive shares): list neighbours, which touch the
// same samples, are then processed under one L2. Every block gets a share however small n is.
__device__ __forceinline__ void block_share(uint64_t n, uint64_t& lo, uint64_t& hi) {
  const uint64_t nb = gridDim.x;
  const uint64_t b = (uint64_t)xcd_remap(blockIdx.x, (int64_t)nb);
  const uint64_t per = (n + nb - 1) / nb;
  lo = b * per < n ? b * per : n;
  hi = lo + per < n ? lo + per : n;
}

// Inclusive prefix sum over the 64 lanes: DPP row shifts within each row of 16 lanes, then the
// totals of the lower rows.
__device__ __forceinline__ uint32_t wave_scan_incl(uint32_t v, int lane) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
  const uint32_t t0 = (uint32_t)__builtin_amdgcn_readlane((int)v, 15);
  const uint32_t t1 = (uint32_t)__builtin_amdgcn_readlane((int)v, 31);
  const uint32_t t2 = (uint32_t)__builtin_amdgcn_readlane((int)v, 47);
  const int row = lane >> 4;
  return v + (row >= 1 ? t0 : 0u) + (row >= 2 ? t1 : 0u) + (row >= 3 ? t2 : 0u);
}

}  // namespace dccommon
