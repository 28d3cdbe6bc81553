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
//   k_qef       one thread per active cell: hermite samples, trilinear normals, 4-cell normal
//               averaging, QEF solve (FP64 plain, or FP32 with FP64 fallback) -> vertex.
//   k_emit      one thread per pencil: walks the face bits along k and writes oriented quads in
//               the CPU backend's order.
//   k_bwd_*     adjoint: per-cell stencil contributions in two phases (k_bwd_cell<1>: rows and
//               per-face normal adjoints; k_bwd_cell<2>: each cell's share of its faces'
//               averaged-normal adjoints), the candidate-sample plane (k_bwd_dilate) and a
//               fixed-order per-sample gather (k_bwd_gather).
//
// Reference semantics (file:line in the reference tree) are cited at each device function.
// Compiled with -ffp-contract=off so that FP64 results match the CPU restatement bit for bit.

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "dc_internal.h"
#include "dc_kernels.h"

namespace {

constexpr int kWave = 64;

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

}  // namespace

#include "dc_cell.h"

namespace {

using namespace dccell;

__device__ __forceinline__ void decode_cell(const DcGridDev& G, uint64_t entry, int& i, int& j,
                                            int& k) {
  const int64_t lw = (int64_t)(entry >> 6) - G.act_off;
  const int64_t row = lw / G.W;
  const int c = (int)(lw - row * G.W);
  i = 64 * c + (int)(entry & 63);
  j = (int)(row % G.cy);
  k = (int)(row / G.cy);
}

// One thread per active cell (persistent grid-stride over the device-side count).
template <int PREC>
__global__ void __launch_bounds__(256)
    k_qef(const DcGridDev* __restrict__ grids, int n_grids, const int64_t* __restrict__ act_bounds,
          const uint64_t* __restrict__ act, const uint64_t* __restrict__ alist,
          DcCounters* __restrict__ ctr, uint64_t cap, double* __restrict__ v64,
          float* __restrict__ v32, uint32_t* __restrict__ flags, const uint64_t* __restrict__ ft,
          const uint32_t* __restrict__ pw, const void* __restrict__ nrm, uint64_t capF) {
  const uint64_t n = min(ctr->n_active, cap);
  const uint64_t nf = min(ctr->n_faces, capF);
  const FaceNormals<double> fn64{ft, pw, reinterpret_cast<const double*>(nrm), nullptr,
                                 PREC == DC_PREC_FP64 ? nf : 0ull};
  const FaceNormals<double> inline64{ft, pw, nullptr, nullptr, 0ull};
  const FaceNormals<float> fn32{ft, pw, nullptr, reinterpret_cast<const float*>(nrm),
                                PREC == DC_PREC_FP32 ? nf : 0ull};
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t first = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // iterate whole waves so the fallback ballot below sees a converged wave
  const uint64_t limit = ((n + stride - 1) / stride) * stride;
  for (uint64_t v = first; v < limit; v += stride) {
    bool fb = false;
    if (v < n) {
      const uint64_t entry = alist[v];
      const int g = find_bucket(act_bounds, n_grids, (int64_t)(entry >> 6));
      const DcGridDev G = grids[g];
      int i, j, k;
      decode_cell(G, entry, i, j, k);
      double x[3];
      if (PREC == DC_PREC_FP64) {
        solve_cell_fp64(G, act, fn64, i, j, k, x);
        v64[3 * v + 0] = x[0];
        v64[3 * v + 1] = x[1];
        v64[3 * v + 2] = x[2];
      } else {
        float xl[3];
        if (solve_cell_fp32(G, act, fn32, i, j, k, xl)) {
          x[0] = coordd(G.ox, i, G.vx) + (double)xl[0];
          x[1] = coordd(G.oy, j, G.vy) + (double)xl[1];
          x[2] = coordd(G.oz, k, G.vz) + (double)xl[2];
        } else {
          solve_cell_fp64(G, act, inline64, i, j, k, x);
          fb = true;
        }
        flags[v] = fb ? 1u : 0u;
      }
      v32[3 * v + 0] = (float)x[0];
      v32[3 * v + 1] = (float)x[1];
      v32[3 * v + 2] = (float)x[2];
    }
    if (PREC == DC_PREC_FP32) {
      const uint64_t nfb = __popcll(__ballot(fb));
      if (nfb && (threadIdx.x & 63) == 0) atomicAdd((unsigned long long*)&ctr->fallback, nfb);
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Adjoint: per-cell stencil contributions, then a deterministic gather per sample.
// ---------------------------------------------------------------------------------------------
constexpr int kBwdThreads = 64;
constexpr int kStencil = 64;  // 4x4x4 samples around a cell: offsets -1..2 per axis

template <int PHASE>
__global__ void __launch_bounds__(kBwdThreads)
    k_bwd_cell(const DcGridDev* __restrict__ grids, int n_grids,
               const int64_t* __restrict__ act_bounds, const uint64_t* __restrict__ act,
               const uint64_t* __restrict__ alist, const uint32_t* __restrict__ flags,
               int precision, const DcCounters* __restrict__ ctr, uint64_t cap,
               const float* __restrict__ gv, const uint64_t* __restrict__ ft,
               const uint32_t* __restrict__ pw, const void* __restrict__ nrm, uint64_t capF,
               float* __restrict__ ubar, float* __restrict__ contrib) {
  __shared__ float s_lc[kBwdThreads][kStencil + 1];
  float* lc = s_lc[threadIdx.x];
  const uint64_t n = min(ctr->n_active, cap);
  const uint64_t nf = min(ctr->n_faces, capF);
  const double* n64 = precision == DC_PREC_FP64 ? reinterpret_cast<const double*>(nrm) : nullptr;
  const float* n32 = precision == DC_PREC_FP64 ? nullptr : reinterpret_cast<const float*>(nrm);
  const FaceNormals<double> fd{ft, pw, n64, n32, nf};
  const FaceNormals<float> ff{ft, pw, n64, n32, nf};
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (uint64_t)gridDim.x * blockDim.x) {
    float* out = contrib + v * kStencil;
    if (PHASE == 1) {
      for (int q = 0; q < kStencil; ++q) lc[q] = 0.f;
    } else {
      for (int q = 0; q < kStencil; ++q) lc[q] = out[q];
    }
    const uint64_t entry = alist[v];
    const int g = find_bucket(act_bounds, n_grids, (int64_t)(entry >> 6));
    const DcGridDev G = grids[g];
    int i, j, k;
    decode_cell(G, entry, i, j, k);
    const bool f64 = precision == DC_PREC_FP64 || flags[v];
    if (PHASE == 1) {
      const float xb[3] = {gv[3 * v], gv[3 * v + 1], gv[3 * v + 2]};
      if (f64)
        backward_rows<double>(G, act, fd, i, j, k, xb, lc, ubar);
      else
        backward_rows<float>(G, act, ff, i, j, k, xb, lc, ubar);
    } else {
      if (f64)
        backward_faces<double>(G, act, fd, i, j, k, ubar, lc);
      else
        backward_faces<float>(G, act, ff, i, j, k, ubar, lc);
    }
    for (int q = 0; q < kStencil; ++q) out[q] = lc[q];
  }
}

// Candidate samples: bit x of word (z*ny + y)*Ws + x/64 is set when some active cell lies in
// [x-2, x+1] x [y-2, y+1] x [z-2, z+1] (the cells whose 4x4x4 stencil contains the sample).
__global__ void __launch_bounds__(256)
    k_bwd_dilate(const DcGridDev* __restrict__ grids, int n_grids,
                 const int64_t* __restrict__ dil_bounds, int64_t n_dil,
                 const uint64_t* __restrict__ act, uint64_t* __restrict__ dil) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_dil) return;
  const int g = find_bucket(dil_bounds, n_grids, t);
  const DcGridDev G = grids[g];
  const int64_t local = t - dil_bounds[g];
  const int64_t row = local / G.Ws;
  const int c = (int)(local - row * G.Ws);
  const int y = (int)(row % G.ny), z = (int)(row / G.ny);
  uint64_t am = 0, a0 = 0, ap = 0;  // OR of chunk c-1, c, c+1 over the 16 cell rows
  for (int kz = max(z - 2, 0); kz <= min(z + 1, G.cz - 1); ++kz)
    for (int jy = max(y - 2, 0); jy <= min(y + 1, G.cy - 1); ++jy) {
      const uint64_t* w = act + G.act_off + ((int64_t)kz * G.cy + jy) * G.W;
      if (c - 1 >= 0 && c - 1 < G.W) am |= w[c - 1];
      if (c < G.W) a0 |= w[c];
      if (c + 1 < G.W) ap |= w[c + 1];
    }
  const uint64_t d = a0 | ((a0 >> 1) | (ap << 63)) | ((a0 << 1) | (am >> 63)) |
                     ((a0 << 2) | (am >> 62));
  dil[G.dil_off + local] = d;
}

__global__ void __launch_bounds__(256)
    k_bwd_gather(const DcGridDev* __restrict__ grids, int n_grids,
                 const int64_t* __restrict__ samp_bounds, int64_t n_samples,
                 const uint64_t* __restrict__ act, const uint32_t* __restrict__ wprefix,
                 const uint64_t* __restrict__ dil, const float* __restrict__ contrib,
                 float* __restrict__ grad) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_samples) return;
  const int g = find_bucket(samp_bounds, n_grids, s);
  const DcGridDev G = grids[g];
  const int64_t local = s - samp_bounds[g];
  const int64_t row = local / G.nx;
  const int x = (int)(local - row * G.nx);
  const int y = (int)(row % G.ny), z = (int)(row / G.ny);
  const uint64_t d = dil[G.dil_off + row * G.Ws + (x >> 6)];
  float acc = 0.f;
  if ((d >> (x & 63)) & 1ull) {
    // fixed order: cells by (k, j, i) ascending
    for (int kz = max(z - 2, 0); kz <= min(z + 1, G.cz - 1); ++kz)
      for (int jy = max(y - 2, 0); jy <= min(y + 1, G.cy - 1); ++jy) {
        const int64_t rw = G.act_off + ((int64_t)kz * G.cy + jy) * G.W;
        for (int ix = max(x - 2, 0); ix <= min(x + 1, G.cx - 1); ++ix) {
          const uint64_t w = act[rw + (ix >> 6)];
          const uint64_t bit = 1ull << (ix & 63);
          if (!(w & bit)) continue;
          const uint64_t vid = (uint64_t)wprefix[rw + (ix >> 6)] + __popcll(w & (bit - 1ull));
          const int slot = (z - kz + 1) * 16 + (y - jy + 1) * 4 + (x - ix + 1);
          acc += contrib[vid * kStencil + slot];
        }
      }
  }
  grad[s] = acc;
}

// ---------------------------------------------------------------------------------------------
// k_emit — quads in the CPU backend's order (contouring.cpp:607-729)
// ---------------------------------------------------------------------------------------------
// contouring.cpp:141-207 gradientAt: central differences with clamped neighbours, normalized.
__device__ __forceinline__ void grid_gradient(const DcGridDev& G, int i, int j, int k, double out[3]) {
  const int im = max(i - 1, 0), ip = min(i + 1, G.nx - 1);
  const int jm = max(j - 1, 0), jp = min(j + 1, G.ny - 1);
  const int km = max(k - 1, 0), kp = min(k + 1, G.nz - 1);
  out[0] = out[1] = out[2] = 0.0;
  const double dx = coordd(G.ox, ip, G.vx) - coordd(G.ox, im, G.vx);
  if (fabs(dx) > 1e-12)
    out[0] = ((double)*sample_ptr(G, ip, j, k) - (double)*sample_ptr(G, im, j, k)) / dx;
  const double dy = coordd(G.oy, jp, G.vy) - coordd(G.oy, jm, G.vy);
  if (fabs(dy) > 1e-12)
    out[1] = ((double)*sample_ptr(G, i, jp, k) - (double)*sample_ptr(G, i, jm, k)) / dy;
  const double dz = coordd(G.oz, kp, G.vz) - coordd(G.oz, km, G.vz);
  if (fabs(dz) > 1e-12)
    out[2] = ((double)*sample_ptr(G, i, j, kp) - (double)*sample_ptr(G, i, j, km)) / dz;
  const double nrm = sqrt(sq3(out[0], out[1], out[2]));
  if (nrm > 0.0) {
    out[0] = out[0] / nrm;
    out[1] = out[1] / nrm;
    out[2] = out[2] / nrm;
  }
}

__device__ __forceinline__ int32_t vertex_id(const DcGridDev& G, const uint64_t* __restrict__ act,
                                             const uint32_t* __restrict__ wprefix, uint32_t vbase,
                                             int i, int j, int k) {
  const int64_t w = G.act_off + ((int64_t)k * G.cy + j) * G.W + (i >> 6);
  const uint64_t below = (1ull << (i & 63)) - 1ull;
  return (int32_t)(wprefix[w] - vbase + (uint32_t)__popcll(act[w] & below));
}

// Faces of the wave's 64 pencils spread over all 64 lanes: lane counts, wave prefix, then each
// lane takes face f = batch + lane, finds its pencil by a branch-free search over the prefixes and
// its k by selecting the r-th set bit of the pencil's face words. fn(G, a, i, j, k, face) runs
// once per face; face is the global face index (pbase of the pencil + rank in the pencil).
template <typename Fn>
__device__ __forceinline__ void for_each_face(const DcGridDev* __restrict__ grids, int n_grids,
                                              const int64_t* __restrict__ pen_bounds,
                                              int64_t n_pen, const uint64_t* __restrict__ ft,
                                              const uint32_t* __restrict__ pbase,
                                              uint32_t* __restrict__ pw, Fn&& fn) {
  const int lane = threadIdx.x & 63;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t cnt = 0;
  if (p < n_pen) {
    const int g = find_bucket(pen_bounds, n_grids, p);
    const DcGridDev& G = grids[g];
    const int64_t local = p - G.pen_off;
    const int64_t plane = (int64_t)G.cy * G.cx;
    const int a = (int)(local / plane);
    const int64_t rem = local - (int64_t)a * plane;
    uint32_t pos = pbase[p];
    for (int kw = 0; kw < G.Wk; ++kw) {
      const int64_t idx = G.ft_off + ((int64_t)a * G.Wk + kw) * plane + rem;
      if (pw) pw[idx] = pos;
      const uint32_t c = (uint32_t)__popcll(ft[idx]);
      pos += c;
      cnt += c;
    }
  }
  uint32_t incl = cnt;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(incl, d);
    if (lane >= d) incl += y;
  }
  const uint32_t off = incl - cnt;
  const uint32_t total = __shfl(incl, 63);
  for (uint32_t batch = 0; batch < total; batch += 64) {
    const uint32_t f = batch + lane;
    int owner = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1) {
      const int mid = owner + step;
      const uint32_t om = __shfl(off, mid & 63);
      if (mid < 64 && om <= f) owner = mid;
    }
    const uint32_t off_o = __shfl(off, owner);
    const int64_t po = __shfl(p, owner);
    if (f >= total) continue;
    uint32_t r = f - off_o;
    const int g = find_bucket(pen_bounds, n_grids, po);
    const DcGridDev& G = grids[g];
    const int64_t local = po - G.pen_off;
    const int64_t plane = (int64_t)G.cy * G.cx;
    const int a = (int)(local / plane);
    const int64_t rem = local - (int64_t)a * plane;
    const int j = (int)(rem / G.cx);
    const int i = (int)(rem - (int64_t)j * G.cx);
    int k = -1;
    for (int kw = 0; kw < G.Wk; ++kw) {
      uint64_t w = ft[G.ft_off + ((int64_t)a * G.Wk + kw) * plane + rem];
      const uint32_t c = (uint32_t)__popcll(w);
      if (r < c) {
        for (uint32_t t = 0; t < r; ++t) w &= w - 1;
        k = 64 * kw + __ffsll((long long)w) - 1;
        break;
      }
      r -= c;
    }
    if (k < 0) continue;
    fn(G, a, i, j, k, (uint64_t)pbase[po] + (f - off_o));
  }
}

// Averaged hermite normal of every face edge, once per face, in face order (plain: FP64 in the
// global frame; FP32 mode: FP32 in each cell's frame). Also writes the per-word face prefix pw.
template <typename T>
__global__ void __launch_bounds__(256)
    k_face_normals(const DcGridDev* __restrict__ grids, int n_grids,
                   const int64_t* __restrict__ pen_bounds, int64_t n_pen,
                   const uint64_t* __restrict__ ft, const uint32_t* __restrict__ pbase,
                   uint32_t* __restrict__ pw, const DcCounters* __restrict__ ctr, uint64_t capF,
                   T* __restrict__ nrm) {
  for_each_face(grids, n_grids, pen_bounds, n_pen, ft, pbase, pw,
                [&](const DcGridDev& G, int a, int i, int j, int k, uint64_t face) {
                  if (face >= capF) return;
                  T n[3];
                  group_sum(G, a, i, j, k, n);
                  const T len = tsqrt(sq3(n[0], n[1], n[2]));
                  normalize3(n);
                  nrm[4 * face + 0] = n[0];
                  nrm[4 * face + 1] = n[1];
                  nrm[4 * face + 2] = n[2];
                  nrm[4 * face + 3] = len;
                });
}

__global__ void __launch_bounds__(256)
    k_emit(const DcGridDev* __restrict__ grids, int n_grids, const int64_t* __restrict__ pen_bounds,
           int64_t n_pen, const uint64_t* __restrict__ act, const uint32_t* __restrict__ wprefix,
           const uint64_t* __restrict__ ft, const uint32_t* __restrict__ pbase,
           const double* __restrict__ v64, const float* __restrict__ v32, int precision,
           const DcCounters* __restrict__ ctr, uint64_t capA, uint64_t capF,
           int32_t* __restrict__ faces) {
  // an overflowed run is repeated by the host; never index vertices past the buffer
  if (ctr->n_active > capA) return;
  for_each_face(grids, n_grids, pen_bounds, n_pen, ft, pbase, nullptr,
                [&](const DcGridDev& G, int a, int i, int j, int k, uint64_t pos) {
    if (pos >= capF) return;
    const uint32_t vbase = wprefix[G.act_off];
    const int64_t vglob = (int64_t)vbase;  // grid's first vertex in the concatenated arrays
    int c0, c1, c2, c3, e1[3];
    if (a == 0) {
      c0 = vertex_id(G, act, wprefix, vbase, i, j - 1, k - 1);
      c1 = vertex_id(G, act, wprefix, vbase, i, j, k - 1);
      c2 = vertex_id(G, act, wprefix, vbase, i, j, k);
      c3 = vertex_id(G, act, wprefix, vbase, i, j - 1, k);
      e1[0] = i + 1; e1[1] = j; e1[2] = k;
    } else if (a == 1) {
      c0 = vertex_id(G, act, wprefix, vbase, i - 1, j, k - 1);
      c1 = vertex_id(G, act, wprefix, vbase, i, j, k - 1);
      c2 = vertex_id(G, act, wprefix, vbase, i, j, k);
      c3 = vertex_id(G, act, wprefix, vbase, i - 1, j, k);
      e1[0] = i; e1[1] = j + 1; e1[2] = k;
    } else {
      c0 = vertex_id(G, act, wprefix, vbase, i - 1, j - 1, k);
      c1 = vertex_id(G, act, wprefix, vbase, i, j - 1, k);
      c2 = vertex_id(G, act, wprefix, vbase, i, j, k);
      c3 = vertex_id(G, act, wprefix, vbase, i - 1, j, k);
      e1[0] = i; e1[1] = j; e1[2] = k + 1;
    }
    double g0[3], g1[3];
    grid_gradient(G, i, j, k, g0);
    grid_gradient(G, e1[0], e1[1], e1[2], g1);
    const double gn0 = g0[0] + g1[0], gn1 = g0[1] + g1[1], gn2 = g0[2] + g1[2];
    double pa[3], pb[3], pc[3], pd[3];
    const int64_t vv[4] = {vglob + c0, vglob + c1, vglob + c2, vglob + c3};
    double* dst[4] = {pa, pb, pc, pd};
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int d = 0; d < 3; ++d)
        dst[q][d] = (precision == DC_PREC_FP64) ? v64[3 * vv[q] + d] : (double)v32[3 * vv[q] + d];
    const double ba0 = pb[0] - pa[0], ba1 = pb[1] - pa[1], ba2 = pb[2] - pa[2];
    const double da0 = pd[0] - pa[0], da1 = pd[1] - pa[1], da2 = pd[2] - pa[2];
    double q0 = ba1 * da2 - ba2 * da1;
    double q1 = ba2 * da0 - ba0 * da2;
    double q2 = ba0 * da1 - ba1 * da0;
    if (sq3(q0, q1, q2) < 1e-24) {
      const double ca0 = pc[0] - pa[0], ca1 = pc[1] - pa[1], ca2 = pc[2] - pa[2];
      q0 = ca1 * da2 - ca2 * da1;
      q1 = ca2 * da0 - ca0 * da2;
      q2 = ca0 * da1 - ca1 * da0;
    }
    if (sq3(gn0, gn1, gn2) > 0 && sq3(q0, q1, q2) > 0) {
      if ((q0 * gn0 + q1 * gn1) + q2 * gn2 < 0.0) {
        const int t = c1;
        c1 = c3;
        c3 = t;
      }
    }
    reinterpret_cast<int4*>(faces)[pos] = make_int4(c0, c1, c2, c3);
  });
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

hipError_t dc_launch_face_normals(const DcFaceNormalArgs& a, hipStream_t s) {
  if (a.n_pen <= 0) return hipSuccess;
  const dim3 grid((unsigned)((a.n_pen + 255) / 256)), block(256);
  if (a.precision == DC_PREC_FP64)
    hipLaunchKernelGGL(k_face_normals<double>, grid, block, 0, s, a.grids, a.n_grids,
                       a.pen_bounds, a.n_pen, a.ft, a.pbase, a.pw, a.ctr, a.capF,
                       reinterpret_cast<double*>(a.nrm));
  else
    hipLaunchKernelGGL(k_face_normals<float>, grid, block, 0, s, a.grids, a.n_grids,
                       a.pen_bounds, a.n_pen, a.ft, a.pbase, a.pw, a.ctr, a.capF,
                       reinterpret_cast<float*>(a.nrm));
  return hipGetLastError();
}

hipError_t dc_launch_qef(const DcQefArgs& a, hipStream_t s) {
  const dim3 grid((unsigned)a.blocks), block(256);
  if (a.precision == DC_PREC_FP64)
    hipLaunchKernelGGL(k_qef<DC_PREC_FP64>, grid, block, 0, s, a.grids, a.n_grids, a.act_bounds,
                       a.act, a.alist, a.ctr_rw, a.capA, a.v64, a.v32, a.flags, a.ft, a.pw,
                       a.nrm, a.capF);
  else
    hipLaunchKernelGGL(k_qef<DC_PREC_FP32>, grid, block, 0, s, a.grids, a.n_grids, a.act_bounds,
                       a.act, a.alist, a.ctr_rw, a.capA, a.v64, a.v32, a.flags, a.ft, a.pw,
                       a.nrm, a.capF);
  return hipGetLastError();
}

hipError_t dc_launch_backward(const DcBwdArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_bwd_cell<1>, dim3((unsigned)a.cell_blocks), dim3(kBwdThreads), 0, s,
                     a.grids, a.n_grids, a.act_bounds, a.act, a.alist, a.flags, a.precision,
                     a.ctr, a.capA, a.grad_vertices, a.ft, a.pw, a.nrm, a.capF, a.ubar, a.contrib);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bwd_cell<2>, dim3((unsigned)a.cell_blocks), dim3(kBwdThreads), 0, s,
                     a.grids, a.n_grids, a.act_bounds, a.act, a.alist, a.flags, a.precision,
                     a.ctr, a.capA, a.grad_vertices, a.ft, a.pw, a.nrm, a.capF, a.ubar, a.contrib);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (a.n_dil > 0)
    hipLaunchKernelGGL(k_bwd_dilate, dim3((unsigned)((a.n_dil + 255) / 256)), dim3(256), 0, s,
                       a.grids, a.n_grids, a.dil_bounds, a.n_dil, a.act, a.dil);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (a.n_samples > 0)
    hipLaunchKernelGGL(k_bwd_gather, dim3((unsigned)((a.n_samples + 255) / 256)), dim3(256), 0, s,
                       a.grids, a.n_grids, a.samp_bounds, a.n_samples, a.act, a.wprefix, a.dil,
                       a.contrib, a.grad_values);
  return hipGetLastError();
}

hipError_t dc_launch_emit(const DcEmitArgs& a, hipStream_t s) {
  if (a.n_pen <= 0) return hipSuccess;
  const int64_t blocks = (a.n_pen + 255) / 256;  // one lane per pencil, faces balanced per wave
  hipLaunchKernelGGL(k_emit, dim3((unsigned)blocks), dim3(256), 0, s, a.grids, a.n_grids,
                     a.pen_bounds, a.n_pen, a.act, a.wprefix, a.ft, a.pbase, a.v64, a.v32, a.precision,
                     a.ctr, a.capA, a.capF, a.faces);
  return hipGetLastError();
}
