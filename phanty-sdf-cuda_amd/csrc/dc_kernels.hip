// dc_kernels.hip — CDNA4 (gfx950) kernels of the dense dual-contouring path.
//
// Pipeline per batch of grids (a single extraction is a batch of one):
//   k_mark      every sample read once (bit-sliced: lane = sample row, 64-bit masks over x):
//               active-cell plane and the per-axis face words over x.
//   k_ft_transpose  face words over x -> face planes over k (64x64 bit transposes across a
//               wave) and the pencil face counts.
//   scans       per-tile reduce + one-block tile scan + apply, over
//               the active plane (k_scan_act: per-word vertex-id prefix, active-cell list in
//               vertex order), the pencil face counts in the CPU backend's face order
//               (k_scan_pen: first face of every pencil) and the face planes (k_face_list).
//   (the per-cell and per-face kernels are in dc_cells.hip)
//
// Reference semantics (file:line in the reference tree) are cited at each device function.
// Compiled with -ffp-contract=off so that FP64 results match the CPU restatement bit for bit.

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

#include "dc_common.h"
#include "dc_internal.h"
#include "dc_kernels.h"

namespace {

constexpr int kWave = 64;

using dccommon::find_bucket;
using dccommon::wave_scan_incl;
using dccommon::xcd_remap;

// f(std::integral_constant<int, 0>) ... f(std::integral_constant<int, N-1>), in order.
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Lane L of four registers := four wave-uniform values (v_writelane_b32), the other lanes
// unchanged. The values are usually compare results just written by the vector ALU; the leading
// s_nop gives the wait states a lane write needs after such an SGPR write (the compiler cannot
// see into the asm to insert them).
template <int L>
__device__ __forceinline__ void write_lane4(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d,
                                            uint32_t sa, uint32_t sb, uint32_t sc, uint32_t sd) {
  asm("s_nop 3\n\t"
      "v_writelane_b32 %0, %4, %8\n\t"
      "v_writelane_b32 %1, %5, %8\n\t"
      "v_writelane_b32 %2, %6, %8\n\t"
      "v_writelane_b32 %3, %7, %8"
      : "+v"(a), "+v"(b), "+v"(c), "+v"(d)
      : "s"(sa), "s"(sb), "s"(sc), "s"(sd), "i"(L));
}

// A read of memory that no kernel writes, at a wave-uniform address: through the constant
// address space the compiler can issue it as a scalar load (no vector memory instruction).
__device__ __forceinline__ float load_uniform(gfloat p, int64_t idx) {
  return ((const __attribute__((address_space(4))) float*)p)[idx];
}

// ---------------------------------------------------------------------------------------------
// k_mark: bit-sliced. A wave tile is one 64-cell word along x (c) x 56 cell rows (57 sample
// rows) x KL cell layers. Lane r holds sample row j0 + r of the current layer as 64-bit masks
// over x (negative, positive samples), built from one ballot per row of the coalesced loads, so
// the cell and face logic below is a few dozen 64-bit operations per layer for 3584 cells.
// Cell row j0 + r's neighbour row comes from lane r + 1; the sample at x = 64c + 64 (the last
// cell's right corner) is one per-lane load. Outputs per cell layer: the active word (lane r)
// and the three face words over x (ftx), both stored [k][c][j]; k_act_transpose puts the active
// words in vertex order ([k][j][c]) and k_ft_transpose turns the face words into the face planes
// over k. Reference semantics: Cell.cpp:66-76 has_vertex (some corner < 0 and some > 0);
// contouring.cpp:534-539 a quad per strict sign change on an interior edge.
// ---------------------------------------------------------------------------------------------
// Cell rows per wave tile (57 sample rows in lanes 0..56). Per layer a wave keeps one load per
// sample row, the x = 64c + 64 column and 4 stores in flight: 61 vector memory operations, within
// the 63 the wait counter can track (more would force a full drain every layer).
constexpr int kMarkRows = 56;

template <int KL>
__global__ void __launch_bounds__(256)
    k_mark(const DcGridDev* __restrict__ grids, int n_grids, const int64_t* __restrict__ tile_bounds,
           int64_t n_tiles, uint64_t* __restrict__ act_kcj, uint64_t* __restrict__ ftx,
           int64_t plane_words, DcCounters* __restrict__ ctr) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wpb = blockDim.x / kWave;
  const int64_t blk = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t tile = blk * wpb + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  if (tile >= n_tiles) return;  // uniform per wave
  const int g = find_bucket(tile_bounds, n_grids, tile);
  const DcGridDev G = grids[g];
  int64_t lt = tile - tile_bounds[g];
  const int nJB = (G.cy + kMarkRows - 1) / kMarkRows;
  const int c = (int)(lt % G.W);
  lt /= G.W;
  const int jb = (int)(lt % nJB);
  const int kb = (int)(lt / nJB);
  const int j0 = jb * kMarkRows;
  const int k0 = kb * KL;
  const int k1 = min(k0 + KL, G.cz);
  const int jr = j0 + lane;  // this lane's sample row (and cell row for lanes < kMarkRows)
  const bool rowc = lane < kMarkRows && jr < G.cy;
  const bool rowj1 = rowc && jr >= 1;
  const int nvalid = min(64, G.cx - 64 * c);  // cells of this word
  const uint64_t xm = nvalid >= 64 ? ~0ull : ((1ull << nvalid) - 1ull);
  const uint64_t xm1 = c == 0 ? (xm & ~1ull) : xm;  // cells with i >= 1

  // Samples of layers k0..k1 through a bounded buffer resource: a uniform per-row byte offset
  // (no address arithmetic per lane), and reads past the tile's layers return 0. Samples past
  // nx or rows past ny read neighbouring data or 0; only masked bits depend on them.
  const int64_t lbytes = (int64_t)G.nx * G.ny * 4;
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
      (void*)(G.val + (int64_t)k0 * G.nx * G.ny), (short)0,
      (int)(uint32_t)((int64_t)(k1 - k0 + 1) * lbytes), 0x00020000);
  const int vx = (64 * c + lane) * 4;                           // row loads: lane = x
  const int ve = (jr * G.nx + 64 * c + 64) * 4;                 // x = 64c + 64: lane = row
  auto ld = [&](int voff, int soff) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsrc, voff, soff, 0));
  };
  float badacc = 0.f;  // becomes NaN once a NaN or Inf sample has been read
  // Masks of one sample layer for sample row j0 + lane over x: negative (N) and positive (P)
  // samples from v. All rows of layer sl_next are first loaded into w in one burst, so the whole
  // layer is in flight while v is taken apart. v and w swap roles every layer (no copies).
  auto masks = [&](const float (&v)[kMarkRows + 1], float (&w)[kMarkRows + 1], float ev,
                   float& ew, int sl_next, uint64_t& N, uint64_t& P) {
    const int lsoff = (int)((int64_t)(sl_next - k0) * lbytes);
    static_for<kMarkRows + 1>([&](auto rc) {
      constexpr int r = decltype(rc)::value;
      w[r] = ld(vx, lsoff + (j0 + r) * G.nx * 4);
    });
    ew = ld(ve, lsoff);
    uint32_t nlo = 0, nhi = 0, plo = 0, phi = 0;
    static_for<kMarkRows + 1>([&](auto rc) {
      constexpr int r = decltype(rc)::value;
      const uint64_t mn = __ballot(v[r] < 0.f);
      const uint64_t mp = __ballot(v[r] > 0.f);
      badacc = __builtin_fmaf(v[r], 0.f, badacc);
      write_lane4<r>(nlo, nhi, plo, phi, (uint32_t)mn, (uint32_t)(mn >> 32), (uint32_t)mp,
                     (uint32_t)(mp >> 32));
    });
    N = ((uint64_t)nhi << 32) | nlo;
    P = ((uint64_t)phi << 32) | plo;
    (void)ev;
  };
  // outputs in [k][c][j] order (rows fastest), so each store is contiguous across the lanes
  uint64_t* ftx_x = ftx + G.act_off;
  uint64_t* ftx_y = ftx + plane_words + G.act_off;
  uint64_t* ftx_z = ftx + 2 * plane_words + G.act_off;
  uint64_t* actg = act_kcj + G.act_off;

  uint64_t Np = 0, Pp = 0, CAp = 0, CBp = 0;
  // one sample layer from (v, ev): its masks, the face words of cell layer sl (x and y edges lie
  // in sample layer sl) and the active word and z faces of cell layer sl - 1
  auto step = [&](int sl, const float (&v)[kMarkRows + 1], float ev, float (&w)[kMarkRows + 1],
                  float& ew) {
    uint64_t N, P;
    masks(v, w, ev, ew, min(sl + 1, k1), N, P);  // the last layer reloads itself (unused)
    const uint64_t N1 = (N >> 1) | ((uint64_t)(ev < 0.f) << 63);  // sample x+1
    const uint64_t P1 = (P >> 1) | ((uint64_t)(ev > 0.f) << 63);
    const uint64_t A = N | N1, B = P | P1;                         // x-edge has a - / + end
    const uint64_t Nn = __shfl_down(N, 1), Pn = __shfl_down(P, 1);  // sample row j + 1
    const uint64_t CA = A | __shfl_down(A, 1), CB = B | __shfl_down(B, 1);
    if (sl < k1 && rowc) {
      const int k = sl;
      const int64_t wd = ((int64_t)k * G.W + c) * G.cy + jr;
      ftx_x[wd] = (rowj1 && k >= 1) ? (((N & P1) | (P & N1)) & xm) : 0ull;
      ftx_y[wd] = (k >= 1) ? (((N & Pn) | (P & Nn)) & xm1) : 0ull;
    }
    if (sl > k0 && rowc) {
      const int k = sl - 1;
      const int64_t wd = ((int64_t)k * G.W + c) * G.cy + jr;
      actg[wd] = (CAp | CA) & (CBp | CB) & xm;
      ftx_z[wd] = rowj1 ? (((Np & P) | (Pp & N)) & xm1) : 0ull;
    }
    Np = N;
    Pp = P;
    CAp = CA;
    CBp = CB;
  };
  float va[kMarkRows + 1], vb[kMarkRows + 1], ea, eb;
  static_for<kMarkRows + 1>([&](auto rc) {
    constexpr int r = decltype(rc)::value;
    va[r] = ld(vx, (j0 + r) * G.nx * 4);
  });
  ea = ld(ve, 0);
  for (int sl = k0; sl <= k1; sl += 2) {
    step(sl, va, ea, vb, eb);
    if (sl + 1 > k1) break;
    step(sl + 1, vb, eb, va, ea);
  }
  if (__ballot(badacc != badacc) && lane == 0) atomicOr(&ctr->nonfinite, 1u);
}

// 64x64 bit-matrix transpose across a wave: lane l holds row l (bit b = column b) in, column l
// out (six block-swap steps).
__device__ __forceinline__ uint64_t transpose64(uint64_t x) {
  const int lane = threadIdx.x & 63;
  const uint64_t kLo[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                           0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const int sh = 32 >> q;
    const uint64_t m = kLo[q];
    const uint64_t y = __shfl_xor(x, sh);
    x = (lane & sh) ? ((x & ~m) | ((y & ~m) >> sh)) : ((x & m) | ((y & m) << sh));
  }
  return x;
}

// Face words over x ([a][k][c][j]) -> face planes over k (word (a, k/64, j, i), bit k%64) and
// the pencil face counts in face order (a, i, j). One wave per (a, c, kFtRows rows): per 64
// layers, lane = layer loads the rows' words (contiguous) and one wave transpose per row gives
// lane = cell i.

__global__ void __launch_bounds__(256)
    k_ft_transpose(const DcGridDev* __restrict__ grids, int n_grids,
                   const int64_t* __restrict__ tw_bounds, int64_t n_tw,
                   const uint64_t* __restrict__ ftx, int64_t plane_words,
                   uint64_t* __restrict__ ft, uint32_t* __restrict__ pcount) {
  const int lane = threadIdx.x & 63;
  const int64_t t = (int64_t)blockIdx.x * (blockDim.x >> 6) +
                    __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (t >= n_tw) return;  // uniform per wave
  const int g = find_bucket(tw_bounds, n_grids, t);
  const DcGridDev G = grids[g];
  const int ng = (G.cy + kFtRows - 1) / kFtRows;
  int64_t lt = t - tw_bounds[g];
  const int jg = (int)(lt % ng);
  lt /= ng;
  const int c = (int)(lt % G.W);
  const int a = (int)(lt / G.W);
  const int j0 = jg * kFtRows;
  const int i = 64 * c + lane;
  const uint64_t* src = ftx + a * plane_words + G.act_off;
  uint32_t count[kFtRows];
#pragma unroll
  for (int r = 0; r < kFtRows; ++r) count[r] = 0;
  for (int kw = 0; kw < G.Wk; ++kw) {
    const int k = min(64 * kw + lane, G.cz - 1);
    const uint64_t* p = src + ((int64_t)k * G.W + c) * G.cy + j0;
    uint64_t wx[kFtRows];
#pragma unroll
    for (int r = 0; r < kFtRows; ++r) wx[r] = p[min(r, G.cy - 1 - j0)];
#pragma unroll
    for (int r = 0; r < kFtRows; ++r) {
      const uint64_t wk = transpose64(64 * kw + lane < G.cz ? wx[r] : 0ull);
      const int j = j0 + r;
      if (j < G.cy && i < G.cx) {
        ft[G.ft_off + (((int64_t)a * G.Wk + kw) * G.cy + j) * G.cx + i] = wk;
        count[r] += (uint32_t)__popcll(wk);
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kFtRows; ++r) {
    const int j = j0 + r;
    if (j < G.cy && i < G.cx) pcount[G.pen_off + ((int64_t)a * G.cx + i) * G.cy + j] = count[r];
  }
}

// Active words [k][c][j] -> vertex order [k][j][c]. One wave per (k, 64 rows, 8 words), staged
// through LDS so both the loads (lane = row) and the stores (8 consecutive words per row) are
// contiguous.
__global__ void __launch_bounds__(256)
    k_act_transpose(const DcGridDev* __restrict__ grids, int n_grids,
                    const int64_t* __restrict__ at_bounds, int64_t n_at,
                    const uint64_t* __restrict__ act_kcj, uint64_t* __restrict__ act) {
  __shared__ uint64_t s_t[4][64][9];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int64_t t = (int64_t)blockIdx.x * (blockDim.x >> 6) + __builtin_amdgcn_readfirstlane(wv);
  if (t >= n_at) return;  // uniform per wave
  const int g = find_bucket(at_bounds, n_grids, t);
  const DcGridDev G = grids[g];
  const int nj = (G.cy + 63) / 64, nc = (G.W + 7) / 8;
  int64_t lt = t - at_bounds[g];
  const int cb = (int)(lt % nc);
  lt /= nc;
  const int jb = (int)(lt % nj);
  const int k = (int)(lt / nj);
  const int j0 = 64 * jb, c0 = 8 * cb;
  const int nrows = min(64, G.cy - j0), nw = min(8, G.W - c0);
  const uint64_t* src = act_kcj + G.act_off + (int64_t)k * G.W * G.cy;
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (q < nw && lane < nrows) s_t[wv][lane][q] = src[(int64_t)(c0 + q) * G.cy + j0 + lane];
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  uint64_t* dst = act + G.act_off + ((int64_t)k * G.cy + j0) * G.W + c0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int e = 64 * q + lane;  // (row e / 8, word e % 8)
    const int r = e >> 3, w = e & 7;
    if (r < nrows && w < nw) dst[(int64_t)r * G.W + w] = s_t[wv][r][w];
  }
}

// ---------------------------------------------------------------------------------------------
// Scans: per-tile reduce, one-block scan of the tile sums, then apply (three launches; no block
// ever waits on another, and no counter is shared by all blocks).
// ---------------------------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;
constexpr int kTileScanThreads = 1024;

// Wave w of a block owns items [tile*kScanTile + w*64*kScanItems, +64*kScanItems), one per lane
// per round of 64 (coalesced); this is item 0 of the calling lane.
__device__ __forceinline__ int64_t seg_base(int64_t tile) {
  return tile * kScanTile + (int64_t)(threadIdx.x >> 6) * (64 * kScanItems) + (threadIdx.x & 63);
}

__device__ __forceinline__ uint32_t weight(uint64_t w) { return (uint32_t)__popcll(w); }
__device__ __forceinline__ uint32_t weight(uint32_t v) { return v; }

// Loads are clamped rather than guarded (a guard would put every load in its own branch).
template <typename T>
__device__ __forceinline__ T load_clamped(const T* __restrict__ p, int64_t idx, int64_t n) {
  const T v = p[idx < n ? idx : n - 1];
  return idx < n ? v : T(0);
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const T y = __shfl_up(v, d);
    if (lane >= d) v += y;
  }
  return v;
}

// The scan steps as block bodies over an explicit tile index, so that the three scans' steps can
// also share one launch (k_scan3_*).
template <typename T>
__device__ __forceinline__ void tile_reduce(const T* __restrict__ items, int64_t n,
                                            uint32_t* __restrict__ tile_sum, int64_t tile) {
  __shared__ uint32_t s_w[kScanThreads / 64];
  const int64_t base = seg_base(tile);
  uint32_t cnt = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) cnt += weight(load_clamped(items, base + 64 * q, n));
  cnt = wave_sum_u32(cnt);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) t += s_w[w];
    tile_sum[tile] = t;
  }
}

// One block: exclusive scan of the tile sums (each thread takes a run of consecutive tiles).
__device__ __forceinline__ void tile_scan(const uint32_t* __restrict__ tile_sum, int64_t n_tiles,
                                          uint64_t* __restrict__ tile_base,
                                          uint64_t* __restrict__ total) {
  __shared__ uint64_t s_w[kTileScanThreads / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t per = (n_tiles + kTileScanThreads - 1) / kTileScanThreads;
  const int64_t b = (int64_t)threadIdx.x * per;
  uint64_t sum = 0;
  for (int64_t q = 0; q < per; ++q)
    if (b + q < n_tiles) sum += tile_sum[b + q];
  const uint64_t incl = wave_incl_scan(sum);
  if (lane == 63) s_w[wv] = incl;
  __syncthreads();
  uint64_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kTileScanThreads / 64; ++w) {
    const uint64_t t = s_w[w];
    if (w < wv) before += t;
    all += t;
  }
  uint64_t run = before + incl - sum;
  for (int64_t q = 0; q < per; ++q)
    if (b + q < n_tiles) {
      tile_base[b + q] = run;
      run += tile_sum[b + q];
    }
  if (threadIdx.x == 0) *total = all;
}

// Offset of the calling wave's segment: the tile's base plus the block's earlier waves.
__device__ __forceinline__ uint64_t wave_offset(uint32_t wave_total, uint64_t tile_base) {
  __shared__ uint32_t s_w[kScanThreads / 64];
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) s_w[wv] = wave_total;
  __syncthreads();
  uint64_t before = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / 64; ++w)
    if (w < wv) before += s_w[w];
  return tile_base + before;
}

// Writes the entries word * 64 + bit of the set bits of w, from position pos on.
__device__ __forceinline__ void append_bits(uint64_t w, int64_t word, uint64_t pos, uint64_t cap,
                                            uint64_t* __restrict__ list) {
  while (w) {
    const int b = __ffsll((long long)w) - 1;
    if (pos < cap) list[pos] = (uint64_t)word * 64u + (uint64_t)b;
    ++pos;
    w &= w - 1;
  }
}

// contouring.cpp:509-524: vertex id = rank of the active cell in (k, j, i) order. Writes the
// per-word prefix and the active-cell list in vertex order.
__device__ __forceinline__ void scan_act_tile(const uint64_t* __restrict__ act, int64_t n_words,
                                              const uint64_t* __restrict__ tile_base,
                                              uint32_t* __restrict__ wprefix,
                                              uint64_t* __restrict__ alist, uint64_t cap,
                                              int64_t tile) {
  const int64_t base = seg_base(tile);
  uint64_t w[kScanItems];
  uint32_t cnt = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    w[q] = load_clamped(act, base + 64 * q, n_words);
    cnt += (uint32_t)__popcll(w[q]);
  }
  uint64_t run = wave_offset(wave_sum_u32(cnt), tile_base[tile]);
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    const int64_t idx = base + 64 * q;
    const uint32_t c = (uint32_t)__popcll(w[q]);
    const uint32_t incl = wave_scan_incl(c, threadIdx.x & 63);
    const uint64_t pos = run + incl - c;
    if (idx < n_words) wprefix[idx] = (uint32_t)pos;
    append_bits(w[q], idx, pos, cap, alist);
    run += (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
  }
}

// First face of every pencil: exclusive scan of the pencil face counts, which are stored in the
// CPU backend's face order (grid, axis, i, j; k is inside a pencil).
__device__ __forceinline__ void scan_pen_tile(const uint32_t* __restrict__ pcount, int64_t n_pen,
                                              const uint64_t* __restrict__ tile_base,
                                              uint32_t* __restrict__ pbase, int64_t tile) {
  const int64_t base = seg_base(tile);
  uint32_t v[kScanItems];
  uint32_t cnt = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    v[q] = load_clamped(pcount, base + 64 * q, n_pen);
    cnt += v[q];
  }
  uint64_t run = wave_offset(wave_sum_u32(cnt), tile_base[tile]);
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    const int64_t idx = base + 64 * q;
    const uint32_t incl = wave_scan_incl(v[q], threadIdx.x & 63);
    if (idx < n_pen) pbase[idx] = (uint32_t)(run + incl - v[q]);
    run += (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
  }
}

// Face list: every face once, as face word * 64 + bit, in face-word (memory) order, so list
// neighbours are memory neighbours (consecutive cells along x) rather than the face order's
// k-runs. A face's id is pw[word] + popc(ft[word] & below(bit)).
__device__ __forceinline__ void face_list_tile(const uint64_t* __restrict__ ft, int64_t n_ft,
                                               const uint64_t* __restrict__ tile_base,
                                               uint64_t* __restrict__ flist, uint64_t cap,
                                               int64_t tile) {
  const int64_t base = seg_base(tile);
  uint64_t w[kScanItems];
  uint32_t cnt = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    w[q] = load_clamped(ft, base + 64 * q, n_ft);
    cnt += (uint32_t)__popcll(w[q]);
  }
  uint64_t run = wave_offset(wave_sum_u32(cnt), tile_base[tile]);
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    const uint32_t c = (uint32_t)__popcll(w[q]);
    const uint32_t incl = wave_scan_incl(c, threadIdx.x & 63);
    append_bits(w[q], base + 64 * q, run + incl - c, cap, flist);
    run += (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
  }
}

template <typename T>
__global__ void __launch_bounds__(kScanThreads)
    k_tile_reduce(const T* __restrict__ items, int64_t n, uint32_t* __restrict__ tile_sum) {
  tile_reduce(items, n, tile_sum, blockIdx.x);
}

__global__ void __launch_bounds__(kTileScanThreads)
    k_tile_scan(const uint32_t* __restrict__ tile_sum, int64_t n_tiles,
                uint64_t* __restrict__ tile_base, uint64_t* __restrict__ total) {
  tile_scan(tile_sum, n_tiles, tile_base, total);
}

__global__ void __launch_bounds__(kScanThreads)
    k_scan_act(const uint64_t* __restrict__ act, int64_t n_words,
               const uint64_t* __restrict__ tile_base, uint32_t* __restrict__ wprefix,
               uint64_t* __restrict__ alist, uint64_t cap) {
  scan_act_tile(act, n_words, tile_base, wprefix, alist, cap, blockIdx.x);
}

__global__ void __launch_bounds__(kScanThreads)
    k_scan_pen(const uint32_t* __restrict__ pcount, int64_t n_pen,
               const uint64_t* __restrict__ tile_base, uint32_t* __restrict__ pbase) {
  scan_pen_tile(pcount, n_pen, tile_base, pbase, blockIdx.x);
}

__global__ void __launch_bounds__(kScanThreads)
    k_face_list(const uint64_t* __restrict__ ft, int64_t n_ft,
                const uint64_t* __restrict__ tile_base, uint64_t* __restrict__ flist,
                uint64_t cap) {
  face_list_tile(ft, n_ft, tile_base, flist, cap, blockIdx.x);
}

// The three scans (active words, pencil counts, face words) step by step, each step of all three
// in one launch: block b of a step belongs to the active scan (b < ta), the pencil scan (then
// b < ta + tp) or the face scan. The tile sums and bases hold the three scans' tiles in that
// order; the tile-scan step is one block per scan.
__global__ void __launch_bounds__(kScanThreads) k_scan3_reduce(DcScan3Args a) {
  const int64_t b = blockIdx.x, ta = a.tiles_a, tp = a.tiles_p;
  if (b < ta)
    tile_reduce(a.act, a.n_words, a.tile_sum, b);
  else if (b < ta + tp)
    tile_reduce(a.pcount, a.n_pen, a.tile_sum + ta, b - ta);
  else
    tile_reduce(a.ft, a.n_ft, a.tile_sum + ta + tp, b - ta - tp);
}

__global__ void __launch_bounds__(kTileScanThreads) k_scan3_tiles(DcScan3Args a) {
  const int64_t ta = a.tiles_a, tp = a.tiles_p;
  if (blockIdx.x == 0)
    tile_scan(a.tile_sum, ta, a.tile_base, &a.ctr->n_active);
  else if (blockIdx.x == 1)
    tile_scan(a.tile_sum + ta, tp, a.tile_base + ta, &a.ctr->n_faces);
  else
    tile_scan(a.tile_sum + ta + tp, a.tiles_f, a.tile_base + ta + tp, &a.ctr->n_list);
}

__global__ void __launch_bounds__(kScanThreads) k_scan3_apply(DcScan3Args a) {
  const int64_t b = blockIdx.x, ta = a.tiles_a, tp = a.tiles_p;
  if (b < ta)
    scan_act_tile(a.act, a.n_words, a.tile_base, a.wprefix, a.alist, a.capA, b);
  else if (b < ta + tp)
    scan_pen_tile(a.pcount, a.n_pen, a.tile_base + ta, a.pbase, b - ta);
  else
    face_list_tile(a.ft, a.n_ft, a.tile_base + ta + tp, a.flist, a.capF, b - ta - tp);
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

// First vertex and first face of each grid of a batch: the vertex prefix at its first act word
// and the pencil base at its first pencil (out[2q], out[2q+1]).
__global__ void k_grid_offsets(const int64_t* __restrict__ act_b, const int64_t* __restrict__ pen_b,
                               int n, const uint32_t* __restrict__ wprefix,
                               const uint32_t* __restrict__ pbase, uint32_t* __restrict__ out) {
  const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (q >= n) return;
  out[2 * q] = wprefix[act_b[q]];
  out[2 * q + 1] = pbase[pen_b[q]];
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

hipError_t dc_launch_grid_offsets(const int64_t* act_b, const int64_t* pen_b, int n,
                                  const uint32_t* wprefix, const uint32_t* pbase, uint32_t* out,
                                  hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_grid_offsets, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, act_b,
                     pen_b, n, wprefix, pbase, out);
  return hipGetLastError();
}

int64_t dc_mark_tiles(const DcGridDev& g, int layers) {
  const int64_t nJB = (g.cy + kMarkRows - 1) / kMarkRows;
  const int64_t nKB = (g.cz + layers - 1) / layers;
  return (int64_t)g.W * nJB * nKB;
}

hipError_t dc_launch_mark(const DcMarkArgs& a, hipStream_t s) {
  const int wpb = a.waves_per_block;
  if (wpb != 1 && wpb != 2 && wpb != 4) return hipErrorInvalidValue;
  const int64_t blocks = (a.n_tiles + wpb - 1) / wpb;
  if (blocks <= 0) return hipSuccess;
  const dim3 grid((unsigned)blocks), block((unsigned)(64 * wpb));
#define DC_MARK(KL)                                                                                \
  hipLaunchKernelGGL(k_mark<KL>, grid, block, 0, s, a.grids, a.n_grids, a.tile_bounds, a.n_tiles, \
                     a.act_kcj, a.ftx, a.plane_words, a.ctr)
  switch (a.layers) {
    case 8: DC_MARK(8); break;
    case 16: DC_MARK(16); break;
    case 32: DC_MARK(32); break;
    case 64: DC_MARK(64); break;
    default: return hipErrorInvalidValue;
  }
#undef DC_MARK
  return hipGetLastError();
}

hipError_t dc_launch_mark_finish(const DcMarkArgs& a, hipStream_t s) {
  if (a.n_at > 0)
    hipLaunchKernelGGL(k_act_transpose, dim3((unsigned)((a.n_at + 3) / 4)), dim3(256), 0, s,
                       a.grids, a.n_grids, a.at_bounds, a.n_at, a.act_kcj, a.act);
  if (a.n_tw > 0)
    hipLaunchKernelGGL(k_ft_transpose, dim3((unsigned)((a.n_tw + 3) / 4)), dim3(256), 0, s,
                       a.grids, a.n_grids, a.tw_bounds, a.n_tw, a.ftx, a.plane_words, a.ft,
                       a.pcount);
  return hipGetLastError();
}

int64_t dc_scan_tiles(int64_t n) { return (n + kScanTile - 1) / kScanTile; }

namespace {
template <typename T>
hipError_t reduce_and_scan(const T* items, int64_t n, uint32_t* tile_sum, uint64_t* tile_base,
                           uint64_t* total, hipStream_t s) {
  const int64_t tiles = dc_scan_tiles(n);
  hipLaunchKernelGGL(k_tile_reduce<T>, dim3((unsigned)tiles), dim3(kScanThreads), 0, s, items, n,
                     tile_sum);
  hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(kTileScanThreads), 0, s, tile_sum, tiles,
                     tile_base, total);
  return hipGetLastError();
}
}  // namespace

hipError_t dc_launch_scan_act(const uint64_t* act, int64_t n_words, uint32_t* tile_sum,
                              uint64_t* tile_base, uint32_t* wprefix, uint64_t* alist,
                              uint64_t cap, uint64_t* total, hipStream_t s) {
  if (n_words <= 0) return hipSuccess;
  hipError_t e = reduce_and_scan(act, n_words, tile_sum, tile_base, total, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_scan_act, dim3((unsigned)dc_scan_tiles(n_words)), dim3(kScanThreads), 0, s,
                     act, n_words, tile_base, wprefix, alist, cap);
  return hipGetLastError();
}

hipError_t dc_launch_scan_pen(const uint32_t* pcount, int64_t n_pen, uint32_t* tile_sum,
                              uint64_t* tile_base, uint32_t* pbase, uint64_t* total,
                              hipStream_t s) {
  if (n_pen <= 0) return hipSuccess;
  hipError_t e = reduce_and_scan(pcount, n_pen, tile_sum, tile_base, total, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_scan_pen, dim3((unsigned)dc_scan_tiles(n_pen)), dim3(kScanThreads), 0, s,
                     pcount, n_pen, tile_base, pbase);
  return hipGetLastError();
}

hipError_t dc_launch_scan3(const DcScan3Args& a, hipStream_t s) {
  const int64_t tiles = a.tiles_a + a.tiles_p + a.tiles_f;
  if (tiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scan3_reduce, dim3((unsigned)tiles), dim3(kScanThreads), 0, s, a);
  hipLaunchKernelGGL(k_scan3_tiles, dim3(3), dim3(kTileScanThreads), 0, s, a);
  hipLaunchKernelGGL(k_scan3_apply, dim3((unsigned)tiles), dim3(kScanThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t dc_launch_face_list(const uint64_t* ft, int64_t n_ft, uint32_t* tile_sum,
                               uint64_t* tile_base, uint64_t* flist, uint64_t cap,
                               uint64_t* total, hipStream_t s) {
  if (n_ft <= 0) return hipSuccess;
  hipError_t e = reduce_and_scan(ft, n_ft, tile_sum, tile_base, total, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_face_list, dim3((unsigned)dc_scan_tiles(n_ft)), dim3(kScanThreads), 0, s,
                     ft, n_ft, tile_base, flist, cap);
  return hipGetLastError();
}
