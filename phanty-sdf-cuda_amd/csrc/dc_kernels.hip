// dc_kernels.hip — CDNA4 (gfx950) kernels of the dense dual-contouring path.
//
// Pipeline per batch of grids (a single extraction is a batch of one):
//   k_mark      sign pass: every sample read once; per sample pencil and block of 64 layers,
//               the masks of its negative and positive samples (bits over k).
//   k_cells     cell pass: from the sign masks of each cell's four corner pencils, the face
//               planes over k, the active-cell plane (one 64x64 bit transpose across a wave per
//               word) and the pencil face counts.
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

// n := 2n + (v < 0), p := 2p + (v > 0): a compare into VCC and an add with VCC as carry-in per
// mask (the compiler builds the same from a select, a shift and an or, and holds batches of
// compare results in SGPR pairs).
__device__ __forceinline__ void shift_in_sign(uint32_t& n, uint32_t& p, float v) {
  asm("v_cmp_gt_f32 vcc, 0, %2\n\t"
      "v_addc_co_u32 %0, vcc, %0, %0, vcc\n\t"
      "v_cmp_lt_f32 vcc, 0, %2\n\t"
      "v_addc_co_u32 %1, vcc, %1, %1, vcc"
      : "+v"(n), "+v"(p)
      : "v"(v)
      : "vcc");
}

// ---------------------------------------------------------------------------------------------
// Marking, in two passes.
// k_mark (sign pass) reads every sample exactly once. A wave tile is 64 samples along x (lane =
// x) x R sample rows x one block of 64 sample layers. For each of its R pencils a lane builds
// the 64-bit masks over the block's layers of the negative (N) and positive (P) samples, one
// compare and one shift-in per sample and mask, and stores them as one 16-byte sign entry. There
// is no halo: the neighbours a cell needs (x+1, y+1, z+1) come from other entries, read by
// k_cells.
// k_cells (cell pass): lane = cell i of a 64-cell word, one wave per (word, cell row). Per block
// of 64 cell layers it reads the sign entries of the cell's four corner pencils (and bit 0 of
// the next block's), forms the three face words over k (already the face planes' layout) and the
// active word over k, which one 64x64 bit transpose across the wave turns into the active
// plane's word over x.
// Reference semantics: Cell.cpp:66-76 has_vertex (some corner < 0 and some > 0);
// contouring.cpp:534-539 a quad per strict sign change on an interior edge.
// ---------------------------------------------------------------------------------------------
// Sample loads a sign-pass wave keeps in flight: R rows x (kMarkLoads / R) layers.
constexpr int kMarkLoads = 32;

template <int R>
__global__ void __launch_bounds__(256)
    k_mark(const DcGridDev* __restrict__ grids, int n_grids, const int64_t* __restrict__ tile_bounds,
           int64_t n_tiles, uint4* __restrict__ sg, DcCounters* __restrict__ ctr) {
  constexpr int D = kMarkLoads / R;  // layers in flight
  static_assert(D >= 1 && 32 % D == 0, "layers in flight must divide 32");
  const int lane = threadIdx.x & (kWave - 1);
  const int wpb = blockDim.x / kWave;
  const int64_t blk = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t tile = blk * wpb + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  if (tile >= n_tiles) return;  // uniform per wave
  const int g = find_bucket(tile_bounds, n_grids, tile);
  const DcGridDev G = grids[g];
  int64_t lt = tile - tile_bounds[g];
  const int nJB = (G.ny + R - 1) / R;
  const int c = (int)(lt % G.Ws);
  lt /= G.Ws;
  const int jb = (int)(lt % nJB);
  const int kb = (int)(lt / nJB);
  const int j0 = jb * R;
  const int nl = min(64, G.nz - 64 * kb);  // sample layers of this block
  const int64_t lfloats = (int64_t)G.nx * G.ny;
  const uint32_t lbytes = (uint32_t)(lfloats * 4);  // nx * ny < 2^30 - 64 (checked on the host)
  const float* base = G.val + (int64_t)64 * kb * lfloats;
  // Per-lane byte offsets of the tile's rows within a layer. Rows past ny read the last row (not
  // stored); samples past nx read the next row (bits of pencils no cell uses) or, past the
  // layer's end, 0 from the bounded buffer resource.
  uint32_t voff[R];
#pragma unroll
  for (int r = 0; r < R; ++r)
    voff[r] = ((uint32_t)min(j0 + r, G.ny - 1) * (uint32_t)G.nx + (uint32_t)(64 * c + lane)) * 4u;
  // one buffer resource per layer; layers past the block (t >= nl) have no records: their loads
  // touch no memory and read 0
  auto layer = [&](int t) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)(base + (int64_t)t * lfloats), (short)0,
                                             t < nl ? (int)lbytes : 0, 0x00020000);
  };
  float buf[D][R];
  auto load_layer = [&](int t, float (&v)[R]) {
    const __amdgpu_buffer_rsrc_t rs = layer(t);
#pragma unroll
    for (int r = 0; r < R; ++r)
      v[r] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, (int)voff[r], 0, 0));
  };
  // masks of layers 0..31 (lo) and 32..63 (hi), each shifted in from the bottom: the first
  // layer of a half ends in bit 31 and a bit reverse puts layer t in bit t
  uint32_t nlo[R], nhi[R], plo[R], phi[R];
#pragma unroll
  for (int r = 0; r < R; ++r) nlo[r] = nhi[r] = plo[r] = phi[r] = 0u;
  float badacc = 0.f;  // becomes NaN once a NaN or Inf sample has been read
  auto consume = [&](const float (&v)[R], uint32_t (&nm)[R], uint32_t (&pm)[R]) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      badacc = __builtin_fmaf(v[r], 0.f, badacc);
      shift_in_sign(nm[r], pm[r], v[r]);
    }
  };
  static_for<D>([&](auto dc) { load_layer(decltype(dc)::value, buf[decltype(dc)::value]); });
  // layers 0..31, loading D layers ahead (at most layer 31 + D < 64)
  for (int t0 = 0; t0 < 32; t0 += D) {
    static_for<D>([&](auto dc) {
      constexpr int d = decltype(dc)::value;
      consume(buf[d], nlo, plo);
      load_layer(t0 + d + D, buf[d]);
    });
  }
  // layers 32..63: the last D are consumed without loads
  for (int t0 = 32; t0 < 64 - D; t0 += D) {
    static_for<D>([&](auto dc) {
      constexpr int d = decltype(dc)::value;
      consume(buf[d], nhi, phi);
      load_layer(t0 + d + D, buf[d]);
    });
  }
  static_for<D>([&](auto dc) { consume(buf[decltype(dc)::value], nhi, phi); });
  const int64_t nxp = 64 * (int64_t)G.Ws;
  uint4* out = sg + G.sg_off + ((int64_t)kb * G.ny + j0) * nxp + 64 * c + lane;
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (j0 + r < G.ny)
      out[r * nxp] = make_uint4(__builtin_bitreverse32(nlo[r]), __builtin_bitreverse32(nhi[r]),
                                __builtin_bitreverse32(plo[r]), __builtin_bitreverse32(phi[r]));
  if (__ballot(badacc != badacc) && lane == 0) atomicOr(&ctr->nonfinite, 1u);
}

// The sign pass over whole rows (grids with nx % 8 == 0 and 16-byte aligned rows): lane = 8
// consecutive samples along x, read as two 16-byte loads per row and layer, so a wave's load
// covers 1 KiB of one row; one wave per (512-sample row segment, sample row, block of 64 layers),
// kMarkWideDepth layers in flight. The sign entries are k_mark's, bit for bit (pencils past nx
// are not written: no cell reads them).
constexpr int kMarkWideDepth = 8;

__global__ void __launch_bounds__(256)
    k_mark_wide(const DcGridDev* __restrict__ grids, int n_grids,
                const int64_t* __restrict__ tile_bounds, int64_t n_tiles, uint4* __restrict__ sg,
                DcCounters* __restrict__ ctr) {
  constexpr int D = kMarkWideDepth;
  const int lane = threadIdx.x & (kWave - 1);
  const int wpb = blockDim.x / kWave;
  const int64_t blk = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t tile = blk * wpb + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  if (tile >= n_tiles) return;  // uniform per wave
  const int g = find_bucket(tile_bounds, n_grids, tile);
  const DcGridDev G = grids[g];
  int64_t lt = tile - tile_bounds[g];
  const int nseg = (G.nx + 511) / 512;
  const int xs = (int)(lt % nseg);
  lt /= nseg;
  const int j = (int)(lt % G.ny);
  const int kb = (int)(lt / G.ny);
  const int nl = min(64, G.nz - 64 * kb);  // sample layers of this block
  const int64_t lfloats = (int64_t)G.nx * G.ny;
  const uint32_t lbytes = (uint32_t)(lfloats * 4);  // nx * ny < 2^30 - 64 (checked on the host)
  const float* base = G.val + (int64_t)64 * kb * lfloats;
  const int x0 = 512 * xs + 8 * lane;
  // lanes past the row read the next row (pencils no cell uses) or, past the layer's end, 0
  const uint32_t voff = ((uint32_t)j * (uint32_t)G.nx + (uint32_t)x0) * 4u;
  auto load_layer = [&](int t, float (&v)[8]) {
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(base + (int64_t)t * lfloats), (short)0, t < nl ? (int)lbytes : 0, 0x00020000);
    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    const u4 a = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)voff, 0, 0);
    const u4 b = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)voff + 16, 0, 0);
    v[0] = __uint_as_float(a.x);
    v[1] = __uint_as_float(a.y);
    v[2] = __uint_as_float(a.z);
    v[3] = __uint_as_float(a.w);
    v[4] = __uint_as_float(b.x);
    v[5] = __uint_as_float(b.y);
    v[6] = __uint_as_float(b.z);
    v[7] = __uint_as_float(b.w);
  };
  float buf[D][8];
  uint32_t nlo[8], nhi[8], plo[8], phi[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) nlo[q] = nhi[q] = plo[q] = phi[q] = 0u;
  float badacc = 0.f;  // becomes NaN once a NaN or Inf sample has been read
  auto consume = [&](const float (&v)[8], uint32_t (&nm)[8], uint32_t (&pm)[8]) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      badacc = __builtin_fmaf(v[q], 0.f, badacc);
      shift_in_sign(nm[q], pm[q], v[q]);
    }
  };
  static_for<D>([&](auto dc) { load_layer(decltype(dc)::value, buf[decltype(dc)::value]); });
  for (int t0 = 0; t0 < 32; t0 += D) {
    static_for<D>([&](auto dc) {
      constexpr int d = decltype(dc)::value;
      consume(buf[d], nlo, plo);
      load_layer(t0 + d + D, buf[d]);
    });
  }
  for (int t0 = 32; t0 < 64 - D; t0 += D) {
    static_for<D>([&](auto dc) {
      constexpr int d = decltype(dc)::value;
      consume(buf[d], nhi, phi);
      load_layer(t0 + d + D, buf[d]);
    });
  }
  static_for<D>([&](auto dc) { consume(buf[decltype(dc)::value], nhi, phi); });
  const int64_t nxp = 64 * (int64_t)G.Ws;
  uint4* out = sg + G.sg_off + ((int64_t)kb * G.ny + j) * nxp + x0;
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (x0 + q < G.nx)
      out[q] = make_uint4(__builtin_bitreverse32(nlo[q]), __builtin_bitreverse32(nhi[q]),
                          __builtin_bitreverse32(plo[q]), __builtin_bitreverse32(phi[q]));
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

// Cell pass: sign entries -> face planes over k (word (a, k/64, j, i), bit k%64), the active
// plane in vertex order (word (k, j, c), bit i%64) and the pencil face counts in face order
// (a, i, j). One wave per (word c, cell row j), lane = cell i = 64c + lane, looping over the
// blocks of 64 cell layers; the next block's entries are loaded while the current one is used.
__global__ void __launch_bounds__(256)
    k_cells(const DcGridDev* __restrict__ grids, int n_grids, const int64_t* __restrict__ cw_bounds,
            int64_t n_cw, const uint4* __restrict__ sg, uint64_t* __restrict__ act,
            uint64_t* __restrict__ ft, uint32_t* __restrict__ pcount) {
  const int lane = threadIdx.x & 63;
  // consecutive waves (rows j, j + 1 of one word, whose sign entries overlap) share an L2
  const int64_t t = xcd_remap(blockIdx.x, gridDim.x) * (blockDim.x >> 6) +
                    __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (t >= n_cw) return;  // uniform per wave
  const int g = find_bucket(cw_bounds, n_grids, t);
  const DcGridDev G = grids[g];
  const int64_t lt = t - cw_bounds[g];
  const int c = (int)(lt % G.W);
  const int j = (int)(lt / G.W);
  const int i = 64 * c + lane;
  const uint64_t iv = i < G.cx ? ~0ull : 0ull;          // cell exists
  const uint64_t iv1 = (i >= 1 && i < G.cx) ? ~0ull : 0ull;
  const uint64_t jv1 = j >= 1 ? ~0ull : 0ull;
  const int nxp = 64 * G.Ws;
  const int64_t bstride = (int64_t)G.ny * nxp;  // entries per block of 64 sample layers
  const uint4* s0 = sg + G.sg_off + (int64_t)j * nxp + i;
  const int di = min(i + 1, nxp - 1) - i;  // pencil i+1 (0 past the row's entries: unused lanes)
  // the cell's corner pencils (j, i), (j, i+1), (j+1, i), (j+1, i+1) in one block
  auto load4 = [&](int kb, uint4 (&q)[4]) {
    const uint4* p = s0 + kb * bstride;
    q[0] = p[0];
    q[1] = p[di];
    q[2] = p[nxp];
    q[3] = p[nxp + di];
  };
  auto lo64 = [](const uint4& v) { return ((uint64_t)v.y << 32) | v.x; };  // N
  auto hi64 = [](const uint4& v) { return ((uint64_t)v.w << 32) | v.z; };  // P
  const int64_t astride = (int64_t)G.Wk * G.cy * G.cx;  // face words per axis
  uint64_t* ftg = ft + G.ft_off;
  uint32_t cnt[3] = {0u, 0u, 0u};
  uint4 cur[4], nxt[4];
  load4(0, cur);
  for (int kw = 0; kw < G.Wk; ++kw) {
    if (kw + 1 < G.Wks) {
      load4(kw + 1, nxt);
    } else {  // no sample layer 64 (kw + 1): only cell layers past cz would read it
#pragma unroll
      for (int q = 0; q < 4; ++q) nxt[q] = make_uint4(0u, 0u, 0u, 0u);
    }
    const int kv = min(64, G.cz - 64 * kw);  // cell layers of this block
    const uint64_t km = kv >= 64 ? ~0ull : ((1ull << kv) - 1ull);
    const uint64_t km1 = kw == 0 ? (km & ~1ull) : km;  // layers k >= 1
    const uint64_t N = lo64(cur[0]), P = hi64(cur[0]);
    const uint64_t Nx = lo64(cur[1]), Px = hi64(cur[1]);
    const uint64_t Ny = lo64(cur[2]), Py = hi64(cur[2]);
    // sample layer k + 1 at (j, i): the block shifted down, the next block's bit 0 on top
    const uint64_t Nz = (N >> 1) | ((uint64_t)(nxt[0].x & 1u) << 63);
    const uint64_t Pz = (P >> 1) | ((uint64_t)(nxt[0].z & 1u) << 63);
    const uint64_t fx = ((N & Px) | (P & Nx)) & km1 & jv1 & iv;   // edge (i..i+1, j, k)
    const uint64_t fy = ((N & Py) | (P & Ny)) & km1 & iv1;        // edge (i, j..j+1, k)
    const uint64_t fz = ((N & Pz) | (P & Nz)) & km & jv1 & iv1;   // edge (i, j, k..k+1)
    // some corner < 0 and some > 0 over the 4 pencils and the layers k, k + 1
    const uint64_t A = N | Nx | Ny | lo64(cur[3]);
    const uint64_t B = P | Px | Py | hi64(cur[3]);
    const uint32_t An = nxt[0].x | nxt[1].x | nxt[2].x | nxt[3].x;
    const uint32_t Bn = nxt[0].z | nxt[1].z | nxt[2].z | nxt[3].z;
    const uint64_t Ak = A | (A >> 1) | ((uint64_t)(An & 1u) << 63);
    const uint64_t Bk = B | (B >> 1) | ((uint64_t)(Bn & 1u) << 63);
    const uint64_t av = Ak & Bk & km & iv;
    const int64_t fw = ((int64_t)kw * G.cy + j) * G.cx + i;
    if (iv) {
      ftg[fw] = fx;
      ftg[astride + fw] = fy;
      ftg[2 * astride + fw] = fz;
    }
    cnt[0] += (uint32_t)__popcll(fx);
    cnt[1] += (uint32_t)__popcll(fy);
    cnt[2] += (uint32_t)__popcll(fz);
    const uint64_t aw = transpose64(av);  // lane = cell layer 64 kw + lane, bits over i
    const int k = 64 * kw + lane;
    if (k < G.cz) act[G.act_off + ((int64_t)k * G.cy + j) * G.W + c] = aw;
#pragma unroll
    for (int q = 0; q < 4; ++q) cur[q] = nxt[q];
  }
  if (iv) {
#pragma unroll
    for (int a = 0; a < 3; ++a) pcount[G.pen_off + ((int64_t)a * G.cx + i) * G.cy + j] = cnt[a];
  }
}

// ---------------------------------------------------------------------------------------------
// Scans: per-tile reduce, one-block scan of the tile sums, then apply (three launches; no block
// ever waits on another, and no counter is shared by all blocks).
// ---------------------------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;  // 2048-word tiles: half the tile sums of 4 for the one-block scan
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

int64_t dc_mark_tiles(const DcGridDev& g, int rows) {
  if (rows == kMarkWideRows) return (int64_t)((g.nx + 511) / 512) * g.ny * g.Wks;
  return (int64_t)g.Ws * ((g.ny + rows - 1) / rows) * g.Wks;
}

int64_t dc_cell_waves(const DcGridDev& g) { return (int64_t)g.W * g.cy; }

hipError_t dc_launch_mark(const DcMarkArgs& a, hipStream_t s) {
  const int wpb = a.waves_per_block;
  if (wpb != 1 && wpb != 2 && wpb != 4) return hipErrorInvalidValue;
  const int64_t blocks = (a.n_tiles + wpb - 1) / wpb;
  if (blocks <= 0) return hipSuccess;
  const dim3 grid((unsigned)blocks), block((unsigned)(64 * wpb));
#define DC_MARK(R)                                                                                \
  hipLaunchKernelGGL(k_mark<R>, grid, block, 0, s, a.grids, a.n_grids, a.tile_bounds, a.n_tiles,  \
                     a.sg, a.ctr)
  switch (a.rows) {
    case kMarkWideRows:
      hipLaunchKernelGGL(k_mark_wide, grid, block, 0, s, a.grids, a.n_grids, a.tile_bounds,
                         a.n_tiles, a.sg, a.ctr);
      break;
    case 4: DC_MARK(4); break;
    case 8: DC_MARK(8); break;
    case 16: DC_MARK(16); break;
    case 32: DC_MARK(32); break;
    default: return hipErrorInvalidValue;
  }
#undef DC_MARK
  return hipGetLastError();
}

hipError_t dc_launch_mark_finish(const DcMarkArgs& a, hipStream_t s) {
  if (a.n_cw > 0)
    hipLaunchKernelGGL(k_cells, dim3((unsigned)((a.n_cw + 3) / 4)), dim3(256), 0, s, a.grids,
                       a.n_grids, a.cw_bounds, a.n_cw, a.sg, a.act, a.ft, a.pcount);
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
