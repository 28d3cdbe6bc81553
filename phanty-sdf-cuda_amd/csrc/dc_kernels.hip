// dc_kernels.hip — CDNA4 (gfx950) kernels of the dense dual-contouring path.
//
// Pipeline per batch of grids (a single extraction is a batch of one):
//   k_mark      every sample read once: active-cell bit plane + per-axis face bit planes + pencil
//               face counts. Wave64 tile = 64 cells along x x J cell rows x 64 cell layers; the
//               one-sample halo in x comes from a lane shuffle (lane 63 loads one extra sample),
//               the halo in y/z from the next row/layer already in registers.
//   k_scan_act  decoupled look-back scan over the active plane: per-word vertex-id prefix and the
//               active-cell list in vertex order.
//   k_scan_pen  decoupled look-back scan over pencil face counts in the CPU backend's face
//               order (axis, i, j): first face index of every pencil.
//   (the per-cell and per-face kernels are in dc_cells.hip)
//
// Reference semantics (file:line in the reference tree) are cited at each device function.
// Compiled with -ffp-contract=off so that FP64 results match the CPU restatement bit for bit.

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "dc_common.h"
#include "dc_internal.h"
#include "dc_kernels.h"

namespace {

constexpr int kWave = 64;

using dccommon::find_bucket;
using dccommon::xcd_remap;

// ---------------------------------------------------------------------------------------------
// k_mark
// ---------------------------------------------------------------------------------------------
template <int J>
__global__ void __launch_bounds__(256)
    k_mark(const DcGridDev* __restrict__ grids, int n_grids, const int64_t* __restrict__ tile_bounds,
           int64_t n_tiles, uint64_t* __restrict__ act, uint64_t* __restrict__ ft,
           uint32_t* __restrict__ pcount, DcCounters* __restrict__ ctr) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wpb = blockDim.x / kWave;
  const int64_t blk = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t tile = blk * wpb + (threadIdx.x / kWave);
  if (tile >= n_tiles) return;  // uniform per wave
  const int g = find_bucket(tile_bounds, n_grids, tile);
  const DcGridDev G = grids[g];
  int64_t lt = tile - tile_bounds[g];
  const int nJB = (G.cy + J - 1) / J;
  const int c = (int)(lt % G.W);
  lt /= G.W;
  const int jb = (int)(lt % nJB);
  const int kw = (int)(lt / nJB);
  const int i = 64 * c + lane;
  const int j0 = jb * J;
  const int k0 = kw * 64;
  const int k1 = min(k0 + 64, G.cz);
  const bool lane_in = i < G.nx;
  const int ix = 64 * c + 64;
  const bool extra = (lane == 63) && (ix < G.nx);
  const int64_t rs = G.nx;
  const int64_t ls = (int64_t)G.nx * G.ny;

  uint32_t rows_s = 0, rows_c = 0;
#pragma unroll
  for (int r = 0; r <= J; ++r) rows_s |= (uint32_t)(j0 + r < G.ny) << r;
#pragma unroll
  for (int r = 0; r < J; ++r) rows_c |= (uint32_t)(j0 + r < G.cy) << r;
  const uint32_t rows_j1 = (j0 == 0) ? (rows_c & ~1u) : rows_c;
  const bool cin = i < G.cx;
  const bool cin1 = cin && i >= 1;

  float va[J + 1], ea[J + 1];
  bool bad = false;
  auto fetch = [&](int kk, float* v, float* e) {
    const float* base = G.val + (int64_t)kk * ls + (int64_t)j0 * rs;
#pragma unroll
    for (int r = 0; r <= J; ++r) {
      v[r] = 0.f;
      e[r] = 0.f;
      if (rows_s & (1u << r)) {
        if (lane_in) v[r] = base[(int64_t)r * rs + i];
        if (extra) e[r] = base[(int64_t)r * rs + ix];
      }
    }
  };
  auto bits = [&](const float* v, const float* e, uint32_t& N, uint32_t& P, uint32_t& N1,
                  uint32_t& P1) {
    uint32_t ne = 0, pe = 0;
    N = 0;
    P = 0;
#pragma unroll
    for (int r = 0; r <= J; ++r) {
      N |= (uint32_t)(v[r] < 0.f) << r;
      P |= (uint32_t)(v[r] > 0.f) << r;
      ne |= (uint32_t)(e[r] < 0.f) << r;
      pe |= (uint32_t)(e[r] > 0.f) << r;
      bad |= !(fabsf(v[r]) <= FLT_MAX);
    }
    N1 = __shfl_down(N, 1);
    P1 = __shfl_down(P, 1);
    if (lane == 63) {
      N1 = ne;
      P1 = pe;
    }
  };

  uint32_t Np, Pp, N1p, P1p;
  fetch(k0, va, ea);
  bits(va, ea, Np, Pp, N1p, P1p);
  float vb[J + 1], eb[J + 1];
  fetch(k0 + 1, vb, eb);

  uint64_t fx[J], fy[J], fz[J];
#pragma unroll
  for (int r = 0; r < J; ++r) fx[r] = fy[r] = fz[r] = 0;

  for (int k = k0; k < k1; ++k) {
    // layer k+1 was fetched last iteration; issue layer k+2 now so it is in flight meanwhile
#pragma unroll
    for (int r = 0; r <= J; ++r) {
      va[r] = vb[r];
      ea[r] = eb[r];
    }
    if (k + 2 <= k1) fetch(k + 2, vb, eb);
    uint32_t Nc, Pc, N1c, P1c;
    bits(va, ea, Nc, Pc, N1c, P1c);
    const uint32_t XNp = Np | N1p, XPp = Pp | P1p, XNc = Nc | N1c, XPc = Pc | P1c;
    const uint32_t cn = XNp | (XNp >> 1) | XNc | (XNc >> 1);
    const uint32_t cp = XPp | (XPp >> 1) | XPc | (XPc >> 1);
    // Cell.cpp:66-76 has_vertex; contouring.cpp:534-539 strict sign change for quads
    const uint32_t ab = cin ? (cn & cp & rows_c) : 0u;
    const uint32_t fxb = (cin && k >= 1) ? (((Np & P1p) | (Pp & N1p)) & rows_j1) : 0u;
    const uint32_t fyb = (cin1 && k >= 1) ? (((Np & (Pp >> 1)) | (Pp & (Np >> 1))) & rows_c) : 0u;
    const uint32_t fzb = cin1 ? (((Np & Pc) | (Pp & Nc)) & rows_j1) : 0u;
    const int kk = k - k0;
    uint64_t aw = 0;
#pragma unroll
    for (int r = 0; r < J; ++r) {
      fx[r] |= (uint64_t)((fxb >> r) & 1u) << kk;
      fy[r] |= (uint64_t)((fyb >> r) & 1u) << kk;
      fz[r] |= (uint64_t)((fzb >> r) & 1u) << kk;
      const uint64_t w = __ballot((ab >> r) & 1u);
      if (lane == r) aw = w;
    }
    if (lane < J && ((rows_c >> lane) & 1u))
      act[G.act_off + ((int64_t)k * G.cy + j0 + lane) * G.W + c] = aw;
    Np = Nc;
    Pp = Pc;
    N1p = N1c;
    P1p = P1c;
  }
  if (__any(bad) && lane == 0) atomicOr(&ctr->nonfinite, 1u);
  if (cin) {
#pragma unroll
    for (int r = 0; r < J; ++r) {
      if (!((rows_c >> r) & 1u)) continue;
      const int j = j0 + r;
      const uint64_t w3[3] = {fx[r], fy[r], fz[r]};
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        ft[G.ft_off + (((int64_t)a * G.Wk + kw) * G.cy + j) * G.cx + i] = w3[a];
        const uint32_t n = (uint32_t)__popcll(w3[a]);
        if (n) atomicAdd(&pcount[G.pen_off + ((int64_t)a * G.cy + j) * G.cx + i], n);
      }
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Decoupled look-back scan helpers (single pass; tiles taken in launch order from a counter so
// every tile a block waits on has already started).
// ---------------------------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;
constexpr uint64_t kSpinLimit = 1ull << 22;

__device__ __forceinline__ uint64_t block_exclusive(uint64_t v, uint64_t* s_wave, uint64_t* total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint64_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t y = __shfl_up(x, d);
    if (lane >= d) x += y;
  }
  if (lane == 63) s_wave[wv] = x;
  __syncthreads();
  uint64_t before = 0, all = 0;
  for (int q = 0; q < kScanThreads / 64; ++q) {
    if (q < wv) before += s_wave[q];
    all += s_wave[q];
  }
  *total = all;
  return before + x - v;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
  return v;
}

// Decoupled look-back run by the block's first wave: each probe reads the status of the 64
// nearest predecessors at once, sums aggregates up to the nearest inclusive prefix, and waits
// (bounded) only when a tile before that prefix has not published its aggregate yet.
__device__ uint64_t lookback_wave(uint64_t* __restrict__ status, int64_t tile, uint64_t agg,
                                  DcCounters* __restrict__ ctr) {
  const int lane = threadIdx.x & 63;
  if (tile == 0) {
    if (lane == 0)
      __hip_atomic_store(&status[0], (DC_SCAN_FLAG_PREFIX << 62) | agg, __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_AGENT);
    return 0;
  }
  if (lane == 0)
    __hip_atomic_store(&status[tile], (DC_SCAN_FLAG_AGG << 62) | agg, __ATOMIC_RELEASE,
                       __HIP_MEMORY_SCOPE_AGENT);
  uint64_t excl = 0;
  int64_t base = tile - 1;
  uint64_t spins = 0;
  while (true) {
    const int64_t t = base - lane;
    uint64_t st = DC_SCAN_FLAG_PREFIX << 62;  // before tile 0: an empty prefix
    if (t >= 0) st = __hip_atomic_load(&status[t], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t flag = st >> 62;
    const uint64_t pm = __ballot(flag == DC_SCAN_FLAG_PREFIX);
    const uint64_t zm = __ballot(flag == 0);
    const int fp = pm ? __ffsll((long long)pm) - 1 : 64;
    const int fz = zm ? __ffsll((long long)zm) - 1 : 64;
    if (fz < fp) {
      if (++spins > kSpinLimit) {  // bounded wait: report instead of hanging
        if (lane == 0) atomicOr(&ctr->error, DC_ERR_SCAN_WAIT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    excl += wave_sum_u64(lane <= fp ? (st & DC_SCAN_VALUE_MASK) : 0ull);
    if (fp < 64) break;
    base -= 64;
  }
  if (lane == 0)
    __hip_atomic_store(&status[tile], (DC_SCAN_FLAG_PREFIX << 62) | (excl + agg),
                       __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  return excl;
}

// contouring.cpp:509-524: vertex id = rank of the active cell in (k, j, i) order.
__global__ void __launch_bounds__(kScanThreads)
    k_scan_act(const uint64_t* __restrict__ act, int64_t n_words, uint32_t* __restrict__ wprefix,
               uint64_t* __restrict__ alist, uint64_t cap, uint64_t* __restrict__ status,
               uint32_t* __restrict__ tile_ctr, DcCounters* __restrict__ ctr) {
  __shared__ int64_t s_tile;
  __shared__ uint64_t s_wave[kScanThreads / 64];
  __shared__ uint64_t s_prefix;
  if (threadIdx.x == 0) s_tile = (int64_t)atomicAdd(tile_ctr, 1u);
  __syncthreads();
  const int64_t tile = s_tile;
  const int64_t base = tile * kScanTile + (int64_t)threadIdx.x * kScanItems;
  uint64_t w[kScanItems];
  uint64_t cnt = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    w[q] = (base + q < n_words) ? act[base + q] : 0ull;
    cnt += (uint64_t)__popcll(w[q]);
  }
  uint64_t total;
  const uint64_t excl = block_exclusive(cnt, s_wave, &total);
  if (threadIdx.x < 64) {
    const uint64_t pre = lookback_wave(status, tile, total, ctr);
    if (threadIdx.x == 0) s_prefix = pre;
  }
  __syncthreads();
  uint64_t run = s_prefix + excl;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    if (base + q >= n_words) break;
    wprefix[base + q] = (uint32_t)run;
    uint64_t bitsw = w[q];
    uint64_t rk = run;
    while (bitsw) {
      const int b = __ffsll((long long)bitsw) - 1;
      if (rk < cap) alist[rk] = (uint64_t)(base + q) * 64u + (uint64_t)b;
      ++rk;
      bitsw &= bitsw - 1;
    }
    run = rk;
  }
  if (threadIdx.x == 0 && (tile + 1) * kScanTile >= n_words) ctr->n_active = s_prefix + total;
}

// Pencil face counts in the CPU backend's face order: grid, axis, i, j (k is inside a pencil).
__global__ void __launch_bounds__(kScanThreads)
    k_scan_pen(const DcGridDev* __restrict__ grids, int n_grids,
               const int64_t* __restrict__ pen_bounds, int64_t n_pen,
               const uint32_t* __restrict__ pcount, uint32_t* __restrict__ pbase,
               uint64_t* __restrict__ status, uint32_t* __restrict__ tile_ctr,
               DcCounters* __restrict__ ctr) {
  __shared__ int64_t s_tile;
  __shared__ uint64_t s_wave[kScanThreads / 64];
  __shared__ uint64_t s_prefix;
  if (threadIdx.x == 0) s_tile = (int64_t)atomicAdd(tile_ctr, 1u);
  __syncthreads();
  const int64_t tile = s_tile;
  const int64_t base = tile * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int64_t sidx[kScanItems];
  uint32_t val[kScanItems];
  uint64_t cnt = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    const int64_t e = base + q;
    sidx[q] = -1;
    val[q] = 0;
    if (e < n_pen) {
      const int g = find_bucket(pen_bounds, n_grids, e);
      const int cx = grids[g].cx, cy = grids[g].cy;
      const int64_t local = e - pen_bounds[g];
      const int64_t plane = (int64_t)cx * cy;
      const int64_t a = local / plane;
      const int64_t rem = local - a * plane;
      const int64_t ii = rem / cy, jj = rem - (rem / cy) * cy;
      sidx[q] = pen_bounds[g] + (a * cy + jj) * cx + ii;
      val[q] = pcount[sidx[q]];
      cnt += val[q];
    }
  }
  uint64_t total;
  const uint64_t excl = block_exclusive(cnt, s_wave, &total);
  if (threadIdx.x < 64) {
    const uint64_t pre = lookback_wave(status, tile, total, ctr);
    if (threadIdx.x == 0) s_prefix = pre;
  }
  __syncthreads();
  uint64_t run = s_prefix + excl;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    if (sidx[q] < 0) break;
    pbase[sidx[q]] = (uint32_t)run;
    run += val[q];
  }
  if (threadIdx.x == 0 && (tile + 1) * kScanTile >= n_pen) ctr->n_faces = s_prefix + total;
}

__global__ void k_fill_u64(uint64_t* __restrict__ p, int64_t n, uint64_t v) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x)
    p[q] = v;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x)
    p[q] = v;
}

int64_t fill_blocks(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  return b;
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------------------------
hipError_t dc_launch_fill_u64(uint64_t* p, int64_t n, uint64_t v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)fill_blocks(n)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t dc_launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)fill_blocks(n)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

int64_t dc_mark_tiles(const DcGridDev& g, int rows) {
  const int64_t nJB = (g.cy + rows - 1) / rows;
  return (int64_t)g.W * nJB * g.Wk;
}

hipError_t dc_launch_mark(const DcMarkArgs& a, hipStream_t s) {
  const int wpb = a.waves_per_block;
  if (wpb != 1 && wpb != 2 && wpb != 4) return hipErrorInvalidValue;
  const int64_t blocks = (a.n_tiles + wpb - 1) / wpb;
  if (blocks <= 0) return hipSuccess;
  const dim3 grid((unsigned)blocks), block((unsigned)(64 * wpb));
  switch (a.rows) {
    case 2:
      hipLaunchKernelGGL(k_mark<2>, grid, block, 0, s, a.grids, a.n_grids, a.tile_bounds,
                         a.n_tiles, a.act, a.ft, a.pcount, a.ctr);
      break;
    case 4:
      hipLaunchKernelGGL(k_mark<4>, grid, block, 0, s, a.grids, a.n_grids, a.tile_bounds,
                         a.n_tiles, a.act, a.ft, a.pcount, a.ctr);
      break;
    case 8:
      hipLaunchKernelGGL(k_mark<8>, grid, block, 0, s, a.grids, a.n_grids, a.tile_bounds,
                         a.n_tiles, a.act, a.ft, a.pcount, a.ctr);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

int64_t dc_scan_tiles(int64_t n) { return (n + kScanTile - 1) / kScanTile; }

hipError_t dc_launch_scan_act(const uint64_t* act, int64_t n_words, uint32_t* wprefix,
                              uint64_t* alist, uint64_t cap, uint64_t* status, uint32_t* tile_ctr,
                              DcCounters* ctr, hipStream_t s) {
  const int64_t tiles = dc_scan_tiles(n_words);
  if (tiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scan_act, dim3((unsigned)tiles), dim3(kScanThreads), 0, s, act, n_words,
                     wprefix, alist, cap, status, tile_ctr, ctr);
  return hipGetLastError();
}

hipError_t dc_launch_scan_pen(const DcGridDev* grids, int n_grids, const int64_t* pen_bounds,
                              int64_t n_pen, const uint32_t* pcount, uint32_t* pbase,
                              uint64_t* status, uint32_t* tile_ctr, DcCounters* ctr,
                              hipStream_t s) {
  const int64_t tiles = dc_scan_tiles(n_pen);
  if (tiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scan_pen, dim3((unsigned)tiles), dim3(kScanThreads), 0, s, grids, n_grids,
                     pen_bounds, n_pen, pcount, pbase, status, tile_ctr, ctr);
  return hipGetLastError();
}

