// dc_m2s.hip — mesh to signed distance grid on the GPU: the reference's mesh-to-SDF producer
// (compute_mesh_sdf_device_cuda, src/cuda/mesh_to_sdf_cuda.cu:740-973), with its FP32 arithmetic
// (built without contraction) and its results:
//   bins      triangles listed in the bins (n/8 per axis) their box, grown by 2 max(v), covers
//             (:99-192): (bin, triangle) pairs written triangle-major at scanned offsets, then a
//             stable LSD radix sort by bin, so each bin lists its triangles in ascending order
//             (the reference fills bins with atomics; ascending order is one of its outcomes);
//   evaluate  one workgroup per bin with a triangle in its 3x3x3 neighbourhood: the bin's
//             samples in registers, the neighbourhood's triangles (centre bin first, then the
//             others in (z, y, x) order, :265-271) staged through LDS 256 at a time, every sample
//             taking them in order with the reference's closest-point, tie and sign rules
//             (:226-360). A triangle whose box is farther than the sample's running minimum
//             distance (grown by 1e-4 relative) or than the 2.5-bin band cannot change the
//             sample's value under the 1e-5 tie rule, so it is skipped before the exact test;
//             samples of the other bins are FLT_MAX, written by a streaming fill;
//   flood     far samples as bit masks (64 per word along x); exterior seeds (the border's
//             positive and far samples, every other positive sample, :362-385) flooded through
//             far samples to the fixpoint (:387-414) by tiles of words iterating in LDS until
//             stable, relaunched until no tile changes;
//   relax     4 Jacobi passes (:430-456), the first resolving far samples to +10 (exterior) or
//             -10 (enclosed, :416-428) as it reads them;
//   voxelize  (option) shell voxels within 0.866 voxel diagonals of a triangle (:458-502) as
//             bits, grown by one voxel (:504-527), flooded from the border (:529-569); far
//             samples take -10 in the other (solid) voxels, +10 in flooded ones (:650-679).

#include "dc_m2s.h"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "dc_flagscan.h"

namespace {

using dcscan::exclusive_scan;
using dcscan::kThreads;
using dcscan::scan_tiles;

constexpr float kFar = FLT_MAX;
constexpr int kMaxDim = 4096;                      // samples per axis
constexpr int64_t kMaxTriangles = int64_t(1) << 28;
constexpr int64_t kMaxVertices = int64_t(1) << 31;
constexpr uint64_t kMaxPairs = (uint64_t(1) << 32) - 1;
constexpr int kMaxFloodPasses = 1 << 20;

// counters read back by the host (uint64 slots)
enum { kCtrPairs = 0, kCtrErr, kCtrTris, kCtrMaxAbs, kCtrActive, kCtrFlood, kCtrSlots = 8 };
enum { kErrIndex = 1, kErrNonFinite = 2 };

struct Geom {
  int nx, ny, nz;
  float ox, oy, oz, vx, vy, vz;
  int gx, gy, gz;
  float bdx, bdy, bdz;
  float min_dim, margin, eps_sq, lim, lim_sq_pad, slack;
};

struct V3 {
  float x, y, z;
};
__host__ __device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ __forceinline__ V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ __forceinline__ V3 scale(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__host__ __device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// floor(x) as an index clamped to [0, g - 1]; NaN and out-of-range values clamp (the reference
// clamps after an int conversion, :130-135)
__host__ __device__ __forceinline__ int clamp_index(float x, int g) {
  const float f = floorf(x);
  if (!(f >= 0.0f)) return 0;
  if (f >= (float)(g - 1)) return g - 1;
  return (int)f;
}

// Squared distance from p to triangle (a, b, c) and the closest point: the closest-point regions
// of Ericson, "Real-Time Collision Detection" (2005) 5.1.5, in the reference's operation order
// (:40-97), so every FP32 result is the reference's.
__device__ __forceinline__ float tri_dist_sq(V3 p, V3 a, V3 b, V3 c, V3& q) {
  const V3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
  const float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) {
    q = a;
    return dot(ap, ap);
  }
  const V3 bp = sub(p, b);
  const float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) {
    q = b;
    return dot(bp, bp);
  }
  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    const float t = d1 / (d1 - d3);
    q = add(a, scale(t, ab));
    const V3 r = sub(p, q);
    return dot(r, r);
  }
  const V3 cp = sub(p, c);
  const float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) {
    q = c;
    return dot(cp, cp);
  }
  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    const float t = d2 / (d2 - d6);
    q = add(a, scale(t, ac));
    const V3 r = sub(p, q);
    return dot(r, r);
  }
  const float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
    const float t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    q = add(b, scale(t, sub(c, b)));
    const V3 r = sub(p, q);
    return dot(r, r);
  }
  const float inv = 1.0f / (va + vb + vc);
  const float s = vb * inv, t = vc * inv;
  q = add(add(a, scale(s, ab)), scale(t, ac));
  const V3 r = sub(p, q);
  return dot(r, r);
}

// The angle weight of corner k (:316-339): acos of the clamped cosine between the corner's two
// edges, evaluated in FP64 and rounded to FP32; -1 when an edge is too short (the area weight
// stays).
__device__ __forceinline__ float corner_angle(V3 v0, V3 v1, V3 v2, int k) {
  V3 e1, e2;
  if (k == 0) {
    e1 = sub(v1, v0);
    e2 = sub(v2, v0);
  } else if (k == 1) {
    e1 = sub(v0, v1);
    e2 = sub(v2, v1);
  } else {
    e1 = sub(v0, v2);
    e2 = sub(v1, v2);
  }
  const float l1 = sqrtf(dot(e1, e1)), l2 = sqrtf(dot(e2, e2));
  if (!(l1 > 1e-6f && l2 > 1e-6f)) return -1.0f;
  const float c = fmaxf(-1.0f, fminf(1.0f, dot(e1, e2) / (l1 * l2)));
  return (float)acos((double)c);
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_u64(unsigned long long* __restrict__ p, int64_t n, unsigned long long v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---------------------------------------------------------------------------------------------
// Triangles and bins
// ---------------------------------------------------------------------------------------------
// Per triangle: index and finiteness checks; the evaluation record (corners, unit normal, twice
// the area, corner angles, box: six float4s); the bin box and its bin count (0 for a degenerate
// triangle, whose normal the reference cannot form and skips, :296-302).
__global__ void __launch_bounds__(256)
    k_tris(const float* __restrict__ V, int64_t nv, const int32_t* __restrict__ T, int64_t nt,
           Geom g, float4* __restrict__ rec, uint4* __restrict__ box, uint32_t* __restrict__ cnt,
           unsigned long long* __restrict__ ctr) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nt;
       t += (int64_t)gridDim.x * blockDim.x) {
    cnt[t] = 0u;
    const int32_t i0 = T[3 * t], i1 = T[3 * t + 1], i2 = T[3 * t + 2];
    if (i0 < 0 || i1 < 0 || i2 < 0 || i0 >= nv || i1 >= nv || i2 >= nv) {
      atomicOr(ctr + kCtrErr, (unsigned long long)kErrIndex);
      continue;
    }
    const V3 v0 = {V[3 * (int64_t)i0], V[3 * (int64_t)i0 + 1], V[3 * (int64_t)i0 + 2]};
    const V3 v1 = {V[3 * (int64_t)i1], V[3 * (int64_t)i1 + 1], V[3 * (int64_t)i1 + 2]};
    const V3 v2 = {V[3 * (int64_t)i2], V[3 * (int64_t)i2 + 1], V[3 * (int64_t)i2 + 2]};
    const float m = fmaxf(fmaxf(fmaxf(fabsf(v0.x), fabsf(v0.y)), fmaxf(fabsf(v0.z), fabsf(v1.x))),
                          fmaxf(fmaxf(fabsf(v1.y), fabsf(v1.z)),
                                fmaxf(fmaxf(fabsf(v2.x), fabsf(v2.y)), fabsf(v2.z))));
    // (fmaxf drops NaN operands: each coordinate is tested itself)
    const bool finite = isfinite(v0.x) && isfinite(v0.y) && isfinite(v0.z) && isfinite(v1.x) &&
                        isfinite(v1.y) && isfinite(v1.z) && isfinite(v2.x) && isfinite(v2.y) &&
                        isfinite(v2.z);
    if (!finite) {
      atomicOr(ctr + kCtrErr, (unsigned long long)kErrNonFinite);
      continue;
    }
    atomicMax(ctr + kCtrMaxAbs, (unsigned long long)__float_as_uint(m));
    const V3 n = cross(sub(v1, v0), sub(v2, v0));
    const float len = sqrtf(dot(n, n));
    if (!(len > 1e-12f)) continue;
    const V3 u = {n.x / len, n.y / len, n.z / len};
    const V3 lo = {fminf(v0.x, fminf(v1.x, v2.x)), fminf(v0.y, fminf(v1.y, v2.y)),
                   fminf(v0.z, fminf(v1.z, v2.z))};
    const V3 hi = {fmaxf(v0.x, fmaxf(v1.x, v2.x)), fmaxf(v0.y, fmaxf(v1.y, v2.y)),
                   fmaxf(v0.z, fmaxf(v1.z, v2.z))};
    float4* r = rec + 6 * t;
    r[0] = make_float4(v0.x, v0.y, v0.z, v1.x);
    r[1] = make_float4(v1.y, v1.z, v2.x, v2.y);
    r[2] = make_float4(v2.z, u.x, u.y, u.z);
    r[3] = make_float4(len, corner_angle(v0, v1, v2, 0), corner_angle(v0, v1, v2, 1),
                       corner_angle(v0, v1, v2, 2));
    r[4] = make_float4(lo.x, lo.y, lo.z, hi.x);
    r[5] = make_float4(hi.y, hi.z, 0.0f, 0.0f);
    const int x0 = clamp_index(((lo.x - g.margin) - g.ox) / g.bdx, g.gx);
    const int x1 = clamp_index(((hi.x + g.margin) - g.ox) / g.bdx, g.gx);
    const int y0 = clamp_index(((lo.y - g.margin) - g.oy) / g.bdy, g.gy);
    const int y1 = clamp_index(((hi.y + g.margin) - g.oy) / g.bdy, g.gy);
    const int z0 = clamp_index(((lo.z - g.margin) - g.oz) / g.bdz, g.gz);
    const int z1 = clamp_index(((hi.z + g.margin) - g.oz) / g.bdz, g.gz);
    box[t] = make_uint4((uint32_t)x0 | ((uint32_t)x1 << 16), (uint32_t)y0 | ((uint32_t)y1 << 16),
                        (uint32_t)z0 | ((uint32_t)z1 << 16), 0u);
    const uint32_t c = (uint32_t)(x1 - x0 + 1) * (uint32_t)(y1 - y0 + 1) * (uint32_t)(z1 - z0 + 1);
    cnt[t] = c;
    atomicAdd(ctr + kCtrPairs, (unsigned long long)c);
    atomicAdd(ctr + kCtrTris, 1ull);
  }
}

// One thread per (bin, triangle) pair: the triangle owning pair p (binary search of the scanned
// counts) and the bin of its box at local index p - off[t] (x fastest).
__global__ void __launch_bounds__(256)
    k_pairs(const uint32_t* __restrict__ off, int64_t nt, const uint4* __restrict__ box,
            uint32_t total, int gx, int gy, uint32_t* __restrict__ key, uint32_t* __restrict__ val) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < total;
       p += gridDim.x * blockDim.x) {
    int64_t lo = 0, hi = nt - 1;  // last t with off[t] <= p
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (off[mid] <= p)
        lo = mid;
      else
        hi = mid - 1;
    }
    const uint4 b = box[lo];
    const uint32_t x0 = b.x & 0xFFFFu, x1 = b.x >> 16, y0 = b.y & 0xFFFFu, y1 = b.y >> 16,
                   z0 = b.z & 0xFFFFu;
    const uint32_t lx = x1 - x0 + 1, ly = y1 - y0 + 1;
    const uint32_t l = p - off[lo];
    const uint32_t x = x0 + l % lx, y = y0 + (l / lx) % ly, z = z0 + l / (lx * ly);
    key[p] = x + (uint32_t)gx * (y + (uint32_t)gy * z);
    val[p] = (uint32_t)lo;
  }
}

// Stable LSD radix sort of (key, value) pairs, 8 key bits per pass. Tiles of 2048 pairs per
// block (8 rounds of 256): per-tile digit counts, an exclusive scan of them digit-major, then a
// scatter in which each pair's place among the tile's pairs of its digit is its rank in input
// order (within a wave by digit-matched ballots, across the block's waves by per-wave counts).
constexpr int kSortRounds = 8;
constexpr uint32_t kSortTile = 256 * kSortRounds;

__global__ void __launch_bounds__(256)
    k_sort_hist(const uint32_t* __restrict__ key, uint32_t n, int shift, uint32_t ntiles,
                uint32_t* __restrict__ hist) {
  __shared__ uint32_t c[256];
  c[threadIdx.x] = 0u;
  __syncthreads();
  const uint32_t base = blockIdx.x * kSortTile;
  for (int r = 0; r < kSortRounds; ++r) {
    const uint32_t i = base + r * 256u + threadIdx.x;
    if (i < n) atomicAdd(&c[(key[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = c[threadIdx.x];
}

__global__ void __launch_bounds__(256)
    k_sort_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t n,
                   int shift, uint32_t ntiles, const uint32_t* __restrict__ start,
                   uint32_t* __restrict__ kout, uint32_t* __restrict__ vout) {
  __shared__ uint32_t run[256];
  __shared__ uint32_t wc[4][256];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  run[t] = start[(size_t)t * ntiles + blockIdx.x];
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const uint32_t base = blockIdx.x * kSortTile;
  for (int r = 0; r < kSortRounds; ++r) {
    const uint32_t i = base + r * 256u + t;
    const bool valid = i < n;
    const uint32_t k = valid ? kin[i] : 0u;
    const uint32_t v = valid ? vin[i] : 0u;
    const uint32_t d = (k >> shift) & 255u;
#pragma unroll
    for (int q = 0; q < 4; ++q) wc[q][t] = 0u;
    __syncthreads();
    uint64_t same = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(valid && bit);
      same &= bit ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(same & below);
    if (valid && rank == 0u) wc[w][d] = (uint32_t)__popcll(same);
    __syncthreads();
    if (valid) {
      uint32_t pos = run[d] + rank;
      for (int q = 0; q < w; ++q) pos += wc[q][d];
      kout[pos] = k;
      vout[pos] = v;
    }
    __syncthreads();
    run[t] += wc[0][t] + wc[1][t] + wc[2][t] + wc[3][t];
  }
}

__global__ void k_bin_count(const uint32_t* __restrict__ key, uint32_t n,
                            uint32_t* __restrict__ cnt) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x)
    atomicAdd(cnt + key[p], 1u);
}

// A bin is evaluated when a bin of its 3x3x3 neighbourhood lists a triangle (:194-224).
__global__ void k_bin_active(const uint32_t* __restrict__ cnt, int gx, int gy, int gz,
                             uint32_t* __restrict__ act) {
  const int64_t nb = (int64_t)gx * gy * gz;
  for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < nb;
       b += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(b % gx), y = (int)((b / gx) % gy), z = (int)(b / ((int64_t)gx * gy));
    uint32_t any = 0u;
    for (int zz = max(0, z - 1); zz <= min(gz - 1, z + 1); ++zz)
      for (int yy = max(0, y - 1); yy <= min(gy - 1, y + 1); ++yy)
        for (int xx = max(0, x - 1); xx <= min(gx - 1, x + 1); ++xx)
          any |= cnt[xx + (int64_t)gx * (yy + (int64_t)gy * zz)];
    act[b] = any ? 1u : 0u;
  }
}

__global__ void k_compact(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ idx,
                          int64_t n, uint32_t* __restrict__ list) {
  for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < n;
       b += (int64_t)gridDim.x * blockDim.x)
    if (flag[b]) list[idx[b]] = (uint32_t)b;
}

// ---------------------------------------------------------------------------------------------
// Evaluation
// ---------------------------------------------------------------------------------------------
// Per-axis tables (int32): the bin of each sample coordinate, each bin's first sample (g + 1
// entries), and (voxelize) the voxel of each sample coordinate.
struct Axis {
  const int32_t *bx, *by, *bz, *sx, *sy, *sz, *vx, *vy, *vz;
};

// Samples of the grid whose bin is not evaluated: FLT_MAX.
__global__ void __launch_bounds__(256)
    k_fill_far(Geom g, Axis ax, const uint32_t* __restrict__ act, float* __restrict__ out) {
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
  if (x >= g.nx) return;
  const int64_t b = ax.bx[x] + (int64_t)g.gx * (ax.by[y] + (int64_t)g.gy * ax.bz[z]);
  if (!act[b]) out[x + (int64_t)g.nx * (y + (int64_t)g.ny * z)] = kFar;
}

// The box test. A triangle can matter to a sample only when its computed squared distance d is
// within thr (the sample's running bound, or the band) : d <= thr. Its box [lo, hi] lies farther
// than that from every point of the sample box [a, b] when the squared box distance b2 exceeds
// lim_sq(thr) = ((sqrt(thr) + slack) * 1.0001)^2: slack (1e-5 of the largest coordinate) covers
// the rounding of d, the factor that of b2 and of the hardware square root (no square root per
// test: lim_sq is formed once per bound change).
__device__ __forceinline__ float lim_sq(float thr, float slack) {
  const float r = (__builtin_amdgcn_sqrtf(thr) + slack) * 1.0001f;
  return r * r;
}
__device__ __forceinline__ bool beyond(float lox, float loy, float loz, float hix, float hiy,
                                       float hiz, V3 a, V3 b, float limsq) {
  const float bx = fmaxf(fmaxf(lox - b.x, a.x - hix), 0.0f);
  const float by = fmaxf(fmaxf(loy - b.y, a.y - hiy), 0.0f);
  const float bz = fmaxf(fmaxf(loz - b.z, a.z - hiz), 0.0f);
  return bx * bx + by * by + bz * bz > limsq;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
  return v;
}
// LDS writes of a wave visible to its other lanes
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kEvalPer = 2;  // samples per thread and pass (a bin of 8^3 samples in one pass)

struct SampleState {
  V3 p, q, n;
  float best, bound, limsq;  // bound: running bound on the final minimum; limsq: its box limit
};

// One triangle for one sample (:282-347), after the box test.
__device__ __forceinline__ void visit(SampleState& s, const float4* __restrict__ r, float eps_sq,
                                      float band, float slack) {
  // the box test: skip a triangle that cannot be within the tie tolerance of the final minimum
  const float4 r4 = r[4], r5 = r[5];
  if (beyond(r4.x, r4.y, r4.z, r4.w, r5.x, r5.y, s.p, s.p, s.limsq)) return;
  const float4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
  const V3 v0 = {r0.x, r0.y, r0.z}, v1 = {r0.w, r1.x, r1.y}, v2 = {r1.z, r1.w, r2.x};
  V3 q;
  const float d = tri_dist_sq(s.p, v0, v1, v2, q);
  if ((d + 1e-12f) * 1.0001f < s.bound) {
    s.bound = (d + 1e-12f) * 1.0001f;
    s.limsq = lim_sq(fminf(s.bound, band), slack);
  }
  float w = r3.x;
  const V3 e0 = sub(q, v0), e1 = sub(q, v1), e2 = sub(q, v2);
  float ang = -1.0f;
  if (dot(e0, e0) < eps_sq)
    ang = r3.y;
  else if (dot(e1, e1) < eps_sq)
    ang = r3.z;
  else if (dot(e2, e2) < eps_sq)
    ang = r3.w;
  if (ang >= 0.0f) w = ang;
  const float tol = 1e-5f * (s.best + 1e-8f);
  if (d < s.best - tol) {
    s.best = d;
    s.q = q;
    s.n = {w * r2.y, w * r2.z, w * r2.w};
  } else if (fabsf(d - s.best) <= tol) {
    s.n = {s.n.x + w * r2.y, s.n.y + w * r2.z, s.n.z + w * r2.w};
  }
}

// Samples of a bin are taken in 4x4x4 tiles, one tile per wave and pass, so the lanes that must
// agree before a triangle is skipped lie close together.
constexpr int kTile = 4;

__global__ void __launch_bounds__(256)
    k_eval(Geom g, Axis ax, const uint32_t* __restrict__ act_list,
           const uint32_t* __restrict__ bstart, const uint32_t* __restrict__ tri_of,
           const float4* __restrict__ rec, float* __restrict__ out) {
  __shared__ float4 s_rec[256 * 6];
  __shared__ uint32_t s_beg[27], s_pre[28];
  __shared__ int s_nseg;
  __shared__ uint32_t s_wcnt[4];
  __shared__ float s_wmax[4];
  __shared__ uint8_t s_list[4][256];  // per wave: the chunk entries its samples may still need
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const uint32_t b = act_list[blockIdx.x];
  const int cx = (int)(b % (uint32_t)g.gx), cy = (int)((b / (uint32_t)g.gx) % (uint32_t)g.gy),
            cz = (int)(b / ((uint32_t)g.gx * (uint32_t)g.gy));
  if (t == 0) {  // the visit order: the centre bin, then the others in (z, y, x) order
    int n = 0;
    uint32_t acc = 0;
    auto add_bin = [&](int x, int y, int z) {
      const uint32_t id = (uint32_t)x + (uint32_t)g.gx * ((uint32_t)y + (uint32_t)g.gy * z);
      const uint32_t s0 = bstart[id], s1 = bstart[id + 1];
      s_beg[n] = s0;
      s_pre[n] = acc;
      acc += s1 - s0;
      ++n;
    };
    add_bin(cx, cy, cz);
    for (int z = max(0, cz - 1); z <= min(g.gz - 1, cz + 1); ++z)
      for (int y = max(0, cy - 1); y <= min(g.gy - 1, cy + 1); ++y)
        for (int x = max(0, cx - 1); x <= min(g.gx - 1, cx + 1); ++x)
          if (x != cx || y != cy || z != cz) add_bin(x, y, z);
    s_pre[n] = acc;
    s_nseg = n;
  }
  __syncthreads();
  const int nseg = s_nseg;
  const uint32_t total = s_pre[nseg];
  const int x0 = ax.sx[cx], y0 = ax.sy[cy], z0 = ax.sz[cz];
  const int lx = ax.sx[cx + 1] - x0, ly = ax.sy[cy + 1] - y0, lz = ax.sz[cz + 1] - z0;
  const int tx = (lx + kTile - 1) / kTile, ty = (ly + kTile - 1) / kTile;
  const int nslots = tx * ty * ((lz + kTile - 1) / kTile) * 64;
  const float band = g.lim_sq_pad;
  for (int base = 0; base < nslots; base += 256 * kEvalPer) {
    SampleState S[kEvalPer];
    int64_t idx[kEvalPer];
    bool valid[kEvalPer];
    V3 lo = {kFar, kFar, kFar}, hi = {-kFar, -kFar, -kFar};  // this pass's samples' box
#pragma unroll
    for (int k = 0; k < kEvalPer; ++k) {
      const int s = base + k * 256 + t;
      const int tile = s >> 6, l = s & 63;
      const int ix = x0 + kTile * (tile % tx) + (l & 3);
      const int iy = y0 + kTile * ((tile / tx) % ty) + ((l >> 2) & 3);
      const int iz = z0 + kTile * (tile / (tx * ty)) + (l >> 4);
      valid[k] = s < nslots && ix < x0 + lx && iy < y0 + ly && iz < z0 + lz;
      idx[k] = ix + (int64_t)g.nx * (iy + (int64_t)g.ny * iz);
      S[k].p = {g.ox + (float)ix * g.vx, g.oy + (float)iy * g.vy, g.oz + (float)iz * g.vz};
      S[k].q = {0.0f, 0.0f, 0.0f};
      S[k].n = {0.0f, 1.0f, 0.0f};
      S[k].best = kFar;
      S[k].bound = kFar;
      S[k].limsq = lim_sq(band, g.slack);
    }
    {  // the pass's sample box: the bin's samples (every pass of a bin covers part of them)
      const int xl = x0 + lx - 1, yl = y0 + ly - 1, zl = z0 + lz - 1;
      lo = {g.ox + (float)x0 * g.vx, g.oy + (float)y0 * g.vy, g.oz + (float)z0 * g.vz};
      hi = {g.ox + (float)xl * g.vx, g.oy + (float)yl * g.vy, g.oz + (float)zl * g.vz};
    }
    // this wave's samples' box (empty when none is valid)
    V3 wlo = {kFar, kFar, kFar}, whi = {-kFar, -kFar, -kFar};
#pragma unroll
    for (int k = 0; k < kEvalPer; ++k)
      if (valid[k]) {
        wlo = {fminf(wlo.x, S[k].p.x), fminf(wlo.y, S[k].p.y), fminf(wlo.z, S[k].p.z)};
        whi = {fmaxf(whi.x, S[k].p.x), fmaxf(whi.y, S[k].p.y), fmaxf(whi.z, S[k].p.z)};
      }
    wlo = {wave_min(wlo.x), wave_min(wlo.y), wave_min(wlo.z)};
    whi = {wave_max(whi.x), wave_max(whi.y), wave_max(whi.z)};
    const bool wave_any = whi.x >= wlo.x;
    // box limits of the largest threshold of the block's samples so far, and of this wave's
    float bmax = lim_sq(band, g.slack), wthr = bmax;
    for (uint32_t c0 = 0; c0 < total; c0 += 256) {
      // the chunk's triangles that some sample of the bin may still need, compacted in order
      const uint32_t e = c0 + t;
      bool keep = false;
      uint32_t tri = 0;
      float4 r4 = make_float4(0.f, 0.f, 0.f, 0.f), r5 = r4;
      if (e < total) {
        int sg = 0;
        for (int q = 1; q < nseg; ++q)
          if (s_pre[q] <= e) sg = q;
        tri = tri_of[s_beg[sg] + (e - s_pre[sg])];
        r4 = rec[6 * (size_t)tri + 4];
        r5 = rec[6 * (size_t)tri + 5];
        keep = !beyond(r4.x, r4.y, r4.z, r4.w, r5.x, r5.y, lo, hi, bmax);
      }
      const uint64_t m = __ballot(keep);
      if (lane == 0) s_wcnt[wv] = (uint32_t)__popcll(m);
      __syncthreads();
      uint32_t pos = (uint32_t)__popcll(m & below), cnt = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (q < wv) pos += s_wcnt[q];
        cnt += s_wcnt[q];
      }
      if (keep) {
        float4* d = s_rec + 6 * pos;
#pragma unroll
        for (int j = 0; j < 4; ++j) d[j] = rec[6 * (size_t)tri + j];
        d[4] = r4;
        d[5] = r5;
      }
      __syncthreads();
      // this wave's list: the staged triangles within reach of its own samples, in order
      uint32_t nw = 0;
      if (wave_any)
        for (uint32_t i0 = 0; i0 < cnt; i0 += 64) {
          const uint32_t i = i0 + lane;
          bool need = false;
          if (i < cnt) {
            const float4 a = s_rec[6 * i + 4], c = s_rec[6 * i + 5];
            need = !beyond(a.x, a.y, a.z, a.w, c.x, c.y, wlo, whi, wthr);
          }
          const uint64_t mw = __ballot(need);
          if (need) s_list[wv][nw + (uint32_t)__popcll(mw & below)] = (uint8_t)i;
          nw += (uint32_t)__popcll(mw);
        }
      wave_lds_sync();
      for (uint32_t j = 0; j < nw; ++j) {
        const float4* r = s_rec + 6 * (uint32_t)s_list[wv][j];
#pragma unroll
        for (int k = 0; k < kEvalPer; ++k)
          if (valid[k]) visit(S[k], r, g.eps_sq, band, g.slack);
      }
      float mine = 0.0f;  // (limsq grows with the bound)
#pragma unroll
      for (int k = 0; k < kEvalPer; ++k)
        if (valid[k]) mine = fmaxf(mine, S[k].limsq);
      mine = wave_max(mine);
      wthr = mine;
      if (lane == 0) s_wmax[wv] = mine;
      __syncthreads();
      bmax = fmaxf(fmaxf(s_wmax[0], s_wmax[1]), fmaxf(s_wmax[2], s_wmax[3]));
    }
#pragma unroll
    for (int k = 0; k < kEvalPer; ++k) {
      if (!valid[k]) continue;
      const SampleState& s = S[k];
      float v = kFar;
      if (s.best != kFar) {  // :349-359
        float dist = sqrtf(s.best);
        V3 n = s.n;
        const float nl = sqrtf(dot(n, n));
        if (nl > 1e-12f) n = {n.x / nl, n.y / nl, n.z / nl};
        if (dot(sub(s.p, s.q), n) < 0.0f) dist = -dist;
        v = (dist > g.lim || dist < -g.lim) ? kFar : dist;
      }
      out[idx[k]] = v;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Flood (bit masks: word w of row (y, z) holds x = 64 w .. 64 w + 63 at bits 0..63)
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t fill_up(uint64_t p, uint64_t a) {
  // every bit of p from each bit of a up to the end of its run of p
  return ((p ^ (p + a)) | a) & p;
}
__device__ __forceinline__ uint64_t fill_runs(uint64_t p, uint64_t a) {
  const uint64_t up = fill_up(p, a);
  const uint64_t dn = __builtin_bitreverse64(fill_up(__builtin_bitreverse64(p),
                                                     __builtin_bitreverse64(a)));
  return up | dn;
}

// Sample masks: P = far samples (FLT_MAX), E = the exterior seeds, one wave per word.
__global__ void __launch_bounds__(256)
    k_sample_masks(const float* __restrict__ vals, int nx, int ny, int nz, int W,
                   uint64_t* __restrict__ P, uint64_t* __restrict__ E) {
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6), y = blockIdx.y, z = blockIdx.z;
  if (w >= W) return;
  const int x = 64 * w + lane;
  bool far = false, seed = false;
  if (x < nx) {
    const float f = vals[x + (int64_t)nx * (y + (int64_t)ny * z)];
    const bool border = x == 0 || x == nx - 1 || y == 0 || y == ny - 1 || z == 0 || z == nz - 1;
    far = f == kFar;
    seed = f > 0.0f && (border || !far);
  }
  const uint64_t pf = __ballot(far), sd = __ballot(seed);
  if (lane == 0) {
    const int64_t o = w + (int64_t)W * (y + (int64_t)ny * z);
    P[o] = pf;
    E[o] = sd;
  }
}

// One tile of words per block: the whole rows (W words) of TY x TZ rows, iterated in LDS until no
// word changes (neighbour tiles' words read once as a fixed halo), then written back; `changed`
// is set when a word of the tile changed. Up to 8 words per thread.
constexpr int kFloodWords = 2048;

__global__ void __launch_bounds__(256)
    k_flood(const uint64_t* __restrict__ P, uint64_t* __restrict__ E, int ny, int nz, int W,
            int TY, int TZ, unsigned long long* __restrict__ changed) {
  extern __shared__ uint64_t sE[];  // (TZ + 2) x (TY + 2) x W
  const int t = threadIdx.x;
  const int y0 = blockIdx.x * TY, z0 = blockIdx.y * TZ;
  const int RY = TY + 2, slots = (TZ + 2) * RY * W;
  for (int q = t; q < slots; q += 256) {
    const int w = q % W, r = q / W, yy = r % RY - 1, zz = r / RY - 1;
    const int y = y0 + yy, z = z0 + zz;
    uint64_t e = 0ull;
    if (y >= 0 && y < ny && z >= 0 && z < nz) e = E[w + (int64_t)W * (y + (int64_t)ny * z)];
    sE[q] = e;
  }
  constexpr int kPer = kFloodWords / 256;
  uint64_t pw[kPer], e0[kPer];
  int slot[kPer];
  const int tile = W * TY * TZ;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int q = t + 256 * k;
    const int w = q % W, r = q / W, yy = r % TY, zz = r / TY;
    const bool in = q < tile && y0 + yy < ny && z0 + zz < nz;
    slot[k] = in ? ((zz + 1) * RY + (yy + 1)) * W + w : -1;
    pw[k] = in ? P[w + (int64_t)W * ((y0 + yy) + (int64_t)ny * (z0 + zz))] : 0ull;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPer; ++k) e0[k] = slot[k] >= 0 ? sE[slot[k]] : 0ull;
  while (true) {
    int ch = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int o = slot[k];
      if (o < 0 || !pw[k]) continue;
      const int w = o % W;
      const uint64_t e = sE[o];
      const uint64_t l = w > 0 ? sE[o - 1] : 0ull, r = w < W - 1 ? sE[o + 1] : 0ull;
      const uint64_t nb = (e << 1) | (e >> 1) | (l >> 63) | (r << 63) | sE[o - W] | sE[o + W] |
                          sE[o - RY * W] | sE[o + RY * W];
      const uint64_t ne = e | fill_runs(pw[k], (nb | e) & pw[k]);
      if (ne != e) {
        sE[o] = ne;
        ch = 1;
      }
    }
    if (!__syncthreads_or(ch)) break;
  }
  int wrote = 0;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int o = slot[k];
    if (o < 0) continue;
    const uint64_t e = sE[o];
    if (e != e0[k]) {
      const int q = t + 256 * k;
      const int w = q % W, r = q / W, yy = r % TY, zz = r / TY;
      E[w + (int64_t)W * ((y0 + yy) + (int64_t)ny * (z0 + zz))] = e;
      wrote = 1;
    }
  }
  if (__syncthreads_or(wrote) && t == 0) atomicOr(changed, 1ull);
}

// One relaxation pass (:430-456); with E given, far samples read as +10 (exterior) or -10
// (enclosed cavity, :416-428). One wave per (64 samples along x, sample row y, run of kRelaxRun
// layers), marching up z with the column's values at z - 1, z and z + 1 in registers (the next
// layer's loaded one step ahead): the x neighbours come from the adjacent lanes (the run's edge
// lanes load theirs) and the y neighbours are two loads per layer, so a sample costs about three
// loads instead of seven (csg8 512^3: 989 -> 747 us per pass; loading the whole run's column and
// neighbour rows up front was slower, 859 us). The 4 waves of a block take adjacent rows.
constexpr int kRelaxRun = 16;

__global__ void __launch_bounds__(256)
    k_relax(const float* __restrict__ src, float* __restrict__ dst, Geom g, float keep,
            const uint64_t* __restrict__ E, int W) {
  const int lane = threadIdx.x & 63;
  const int x = blockIdx.x * 64 + lane;
  const int y = blockIdx.y * 4 + (int)(threadIdx.x >> 6);
  const int z0 = blockIdx.z * kRelaxRun;
  if (y >= g.ny) return;  // uniform per wave
  const int z1 = min(z0 + kRelaxRun, g.nz);
  const bool in = x < g.nx;
  const int xc = in ? x : g.nx - 1;  // lanes past the row load a valid sample and store nothing
  const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
  auto value = [&](int xx, int yy, int zz) {
    float f = src[xx + sy * yy + sz * zz];
    if (E && f == kFar)
      f = ((E[(xx >> 6) + (int64_t)W * (yy + (int64_t)g.ny * zz)] >> (xx & 63)) & 1ull) ? 10.0f
                                                                                      : -10.0f;
    return f;
  };
  float fm = z0 > 0 ? value(xc, y, z0 - 1) : 0.0f;
  float f0 = value(xc, y, z0);
  float fp = z0 + 1 < g.nz ? value(xc, y, z0 + 1) : 0.0f;
  for (int z = z0; z < z1; ++z) {
    const float fn = z + 2 < g.nz ? value(xc, y, z + 2) : 0.0f;  // the next layer's, ahead
    // x neighbours from the adjacent lanes; the run's edge lanes load theirs
    float xl = __shfl_up(f0, 1), xr = __shfl_down(f0, 1);
    if (lane == 0 && x > 0) xl = value(x - 1, y, z);
    if (lane == 63 && x + 1 < g.nx) xr = value(x + 1, y, z);
    const float f = f0;
    if (in && !(fabsf(f) <= keep)) {
      float m = fabsf(f);
      if (x > 0) m = fminf(m, fabsf(xl) + g.vx);
      if (x < g.nx - 1) m = fminf(m, fabsf(xr) + g.vx);
      if (y > 0) m = fminf(m, fabsf(value(x, y - 1, z)) + g.vy);
      if (y < g.ny - 1) m = fminf(m, fabsf(value(x, y + 1, z)) + g.vy);
      if (z > 0) m = fminf(m, fabsf(fm) + g.vz);
      if (z < g.nz - 1) m = fminf(m, fabsf(fp) + g.vz);
      dst[x + sy * y + sz * z] = (f < 0.0f ? -1.0f : 1.0f) * m;
    } else if (in) {
      dst[x + sy * y + sz * z] = f;
    }
    fm = f0;
    f0 = fp;
    fp = fn;
  }
}

// ---------------------------------------------------------------------------------------------
// Voxelize option
// ---------------------------------------------------------------------------------------------
struct VoxGeom {
  int vres, W;
  float ox, oy, oz, sx, sy, sz, reach_sq;
};

// Shell voxels: centres within 0.866 voxel diagonals of a triangle (every triangle, :458-502).
__global__ void __launch_bounds__(256)
    k_vox_raster(const float* __restrict__ V, const int32_t* __restrict__ T, int64_t nt, VoxGeom vg,
                 unsigned long long* __restrict__ shell) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nt;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i0 = T[3 * t], i1 = T[3 * t + 1], i2 = T[3 * t + 2];
    const V3 a = {V[3 * i0], V[3 * i0 + 1], V[3 * i0 + 2]};
    const V3 b = {V[3 * i1], V[3 * i1 + 1], V[3 * i1 + 2]};
    const V3 c = {V[3 * i2], V[3 * i2 + 1], V[3 * i2 + 2]};
    const int x0 = clamp_index((fminf(a.x, fminf(b.x, c.x)) - vg.ox) / vg.sx, vg.vres);
    const int x1 = clamp_index((fmaxf(a.x, fmaxf(b.x, c.x)) - vg.ox) / vg.sx, vg.vres);
    const int y0 = clamp_index((fminf(a.y, fminf(b.y, c.y)) - vg.oy) / vg.sy, vg.vres);
    const int y1 = clamp_index((fmaxf(a.y, fmaxf(b.y, c.y)) - vg.oy) / vg.sy, vg.vres);
    const int z0 = clamp_index((fminf(a.z, fminf(b.z, c.z)) - vg.oz) / vg.sz, vg.vres);
    const int z1 = clamp_index((fmaxf(a.z, fmaxf(b.z, c.z)) - vg.oz) / vg.sz, vg.vres);
    for (int z = z0; z <= z1; ++z)
      for (int y = y0; y <= y1; ++y) {
        unsigned long long* row = shell + (int64_t)vg.W * (y + (int64_t)vg.vres * z);
        uint64_t bits = 0ull;
        int word = x0 >> 6;
        for (int x = x0; x <= x1; ++x) {
          if ((x >> 6) != word) {
            if (bits) atomicOr(row + word, (unsigned long long)bits);
            bits = 0ull;
            word = x >> 6;
          }
          const V3 ctr = {vg.ox + (float)x * vg.sx, vg.oy + (float)y * vg.sy,
                          vg.oz + (float)z * vg.sz};
          V3 q;
          if (tri_dist_sq(ctr, a, b, c, q) <= vg.reach_sq) bits |= 1ull << (x & 63);
        }
        if (bits) atomicOr(row + word, (unsigned long long)bits);
      }
  }
}

// The shell grown by one voxel in every direction (3x3x3, :504-527), then P = the voxels outside
// it and E = those of them on the grid's border (:555-569).
__global__ void __launch_bounds__(256)
    k_vox_masks(const uint64_t* __restrict__ shell, VoxGeom vg, uint64_t* __restrict__ P,
                uint64_t* __restrict__ E) {
  const int W = vg.W, n = vg.vres;
  const int64_t words = (int64_t)W * n * n;
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < words;
       o += (int64_t)gridDim.x * blockDim.x) {
    const int w = (int)(o % W), y = (int)((o / W) % n), z = (int)(o / ((int64_t)W * n));
    uint64_t grown = 0ull;
    for (int zz = max(0, z - 1); zz <= min(n - 1, z + 1); ++zz)
      for (int yy = max(0, y - 1); yy <= min(n - 1, y + 1); ++yy) {
        const uint64_t* row = shell + (int64_t)W * (yy + (int64_t)n * zz);
        const uint64_t s = row[w];
        const uint64_t l = w > 0 ? row[w - 1] : 0ull, r = w < W - 1 ? row[w + 1] : 0ull;
        grown |= s | (s << 1) | (s >> 1) | (l >> 63) | (r << 63);
      }
    const int nbits = min(64, n - 64 * w);
    const uint64_t valid = nbits == 64 ? ~0ull : ((1ull << nbits) - 1ull);
    const uint64_t p = ~grown & valid;
    uint64_t border;
    if (y == 0 || y == n - 1 || z == 0 || z == n - 1)
      border = valid;
    else
      border = (w == 0 ? 1ull : 0ull) | (w == W - 1 ? (1ull << (nbits - 1)) : 0ull);
    P[o] = p;
    E[o] = p & border;
  }
}

// Far samples: -10 in solid voxels (not reached by the flood), +10 in the others (:650-679).
__global__ void __launch_bounds__(256)
    k_vox_apply(Geom g, Axis ax, const uint64_t* __restrict__ E, VoxGeom vg,
                float* __restrict__ out) {
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
  if (x >= g.nx) return;
  const int64_t i = x + (int64_t)g.nx * (y + (int64_t)g.ny * z);
  if (out[i] != kFar) return;
  const int vx = ax.vx[x], vy = ax.vy[y], vz = ax.vz[z];
  const bool ext = (E[(vx >> 6) + (int64_t)vg.W * (vy + (int64_t)vg.vres * vz)] >> (vx & 63)) & 1ull;
  out[i] = ext ? 10.0f : -10.0f;
}

int64_t blocks_for(int64_t n, int cu) {
  return std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)cu * 16));
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// DcM2s
// ---------------------------------------------------------------------------------------------
DcM2s::~DcM2s() {
  for (Buf& b : b_)
    if (b.p) (void)hipFree(b.p);
}

int DcM2s::hip_error(hipError_t e, const char* what) {
  if (err_) *err_ = std::string("mesh_to_sdf: ") + what + ": " + hipGetErrorString(e);
  (void)hipGetLastError();
  return e == hipErrorOutOfMemory ? DC_HIP_ERR_OUT_OF_MEMORY : DC_HIP_ERR_HIP;
}

#define M2S_TRY(expr, what)                           \
  do {                                                \
    const hipError_t e_ = (expr);                     \
    if (e_ != hipSuccess) return hip_error(e_, what); \
  } while (0)
#define M2S_RC(expr)                  \
  do {                                \
    const int rc_ = (expr);           \
    if (rc_ != DC_HIP_OK) return rc_; \
  } while (0)
#define M2S_LAUNCH(kernel, grid, block, lds, ...)                          \
  do {                                                                     \
    hipLaunchKernelGGL(kernel, grid, block, lds, s_, __VA_ARGS__);         \
    M2S_TRY(hipGetLastError(), #kernel);                                   \
  } while (0)

int DcM2s::grow(int which, size_t bytes) {
  Buf& b = b_[which];
  if (bytes == 0) bytes = 256;
  if (b.cap >= bytes) return DC_HIP_OK;
  if (b.p) {
    M2S_TRY(hipStreamSynchronize(s_), "synchronize");  // earlier work may still use it
    M2S_TRY(hipFree(b.p), "free");
    b.p = nullptr;
    b.cap = 0;
  }
  const size_t want = bytes + bytes / 8;
  M2S_TRY(hipMalloc(&b.p, want), "allocate");
  b.cap = want;
  return DC_HIP_OK;
}

template <typename T>
static T* at(void* p) {
  return static_cast<T*>(p);
}

// Floods E through P to the fixpoint: tiles of whole rows, relaunched until no tile changes.
int DcM2s::flood(uint64_t* P, uint64_t* E, int nx, int ny, int nz, int* passes) {
  const int W = (nx + 63) / 64;
  // rows per tile: up to kFloodWords words, roughly square in (y, z)
  const int rows = std::max(1, kFloodWords / W);
  int TY = 1;
  while (TY * 2 <= ny && TY * TY * 2 <= rows) TY *= 2;
  const int TZ = std::max(1, std::min(nz, rows / TY));
  const dim3 grid((unsigned)((ny + TY - 1) / TY), (unsigned)((nz + TZ - 1) / TZ));
  const size_t lds = (size_t)(TZ + 2) * (TY + 2) * W * 8;
  unsigned long long* flag = at<unsigned long long>(b_[kCtr].p) + kCtrFlood;
  int n = 0;
  while (true) {
    M2S_LAUNCH(k_fill_u64, dim3(1), dim3(64), 0, flag, (int64_t)1, 0ull);
    M2S_LAUNCH(k_flood, grid, dim3(256), lds, P, E, ny, nz, W, TY, TZ, flag);
    unsigned long long h = 0;
    M2S_TRY(hipMemcpyAsync(&h, flag, 8, hipMemcpyDeviceToHost, s_), "flood flag");
    M2S_TRY(hipStreamSynchronize(s_), "flood flag");
    ++n;
    if (!h) break;
    if (n >= kMaxFloodPasses) {
      if (err_) *err_ = "mesh_to_sdf: the flood did not settle";
      return DC_HIP_ERR_INTERNAL;
    }
  }
  *passes = n;
  return DC_HIP_OK;
}

int DcM2s::run(const float* vertices, int64_t nv, const int32_t* triangles, int64_t nt,
               const int32_t dims[3], const float origin[3], const float spacing[3],
               bool voxelize_first, int32_t voxel_res, float* out, int cu_count, hipStream_t s,
               dc_hip_m2s_stats* st, std::string* err) {
  const auto t0 = std::chrono::steady_clock::now();
  s_ = s;
  cu_ = cu_count > 0 ? cu_count : 256;
  err_ = err;
  dc_hip_m2s_stats S{};
  auto fail = [&](int code, const char* msg) {
    if (err_) *err_ = std::string("mesh_to_sdf: ") + msg;
    return code;
  };
  const int nx = dims[0], ny = dims[1], nz = dims[2];
  if (nx < 1 || ny < 1 || nz < 1 || nx > kMaxDim || ny > kMaxDim || nz > kMaxDim)
    return fail(DC_HIP_ERR_INVALID_GRID, "nx, ny and nz must be in [1, 4096]");
  for (int a = 0; a < 3; ++a)
    if (!std::isfinite(origin[a]) || !std::isfinite(spacing[a]) || !(spacing[a] > 0.0f))
      return fail(DC_HIP_ERR_INVALID_GRID, "origin must be finite and spacing finite and > 0");
  if (nv < 0 || nt < 0 || nv > kMaxVertices || nt > kMaxTriangles)
    return fail(DC_HIP_ERR_INVALID_ARG, "vertex or triangle count out of range");
  if ((nv > 0 && !vertices) || (nt > 0 && !triangles) || !out)
    return fail(DC_HIP_ERR_INVALID_ARG, "NULL vertices, triangles or output");
  const int64_t N = (int64_t)nx * ny * nz;

  // the grid geometry, in the reference's FP32 operations (:807-813)
  Geom g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.ox = origin[0];
  g.oy = origin[1];
  g.oz = origin[2];
  g.vx = spacing[0];
  g.vy = spacing[1];
  g.vz = spacing[2];
  g.gx = std::max(1, nx / 8);
  g.gy = std::max(1, ny / 8);
  g.gz = std::max(1, nz / 8);
  g.bdx = ((float)nx * g.vx) / (float)g.gx;
  g.bdy = ((float)ny * g.vy) / (float)g.gy;
  g.bdz = ((float)nz * g.vz) / (float)g.gz;
  g.min_dim = std::fmin(g.bdx, std::fmin(g.bdy, g.bdz));
  g.margin = 2.0f * std::fmax(g.vx, std::fmax(g.vy, g.vz));
  g.eps_sq = 1e-6f * g.min_dim * g.min_dim;
  g.lim = 2.5f * g.min_dim;
  g.lim_sq_pad = g.lim * g.lim * 1.0001f;
  g.slack = 0.0f;  // set once the coordinates' magnitude is known
  const int64_t nbins = (int64_t)g.gx * g.gy * g.gz;

  // the output: the caller's device buffer, or ours copied out at the end
  bool dev_out = false;
  {
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, out) == hipSuccess && attr.type == hipMemoryTypeDevice)
      dev_out = true;
    (void)hipGetLastError();
  }
  if (!dev_out) M2S_RC(grow(kVals, (size_t)N * 4));
  float* vals = dev_out ? out : at<float>(b_[kVals].p);

  M2S_RC(grow(kCtr, kCtrSlots * 8));
  unsigned long long* ctr = at<unsigned long long>(b_[kCtr].p);
  M2S_LAUNCH(k_fill_u64, dim3(1), dim3(64), 0, ctr, (int64_t)kCtrSlots, 0ull);
  M2S_RC(grow(kInV, (size_t)nv * 12));
  M2S_RC(grow(kInT, (size_t)nt * 12));
  M2S_RC(grow(kTri, (size_t)nt * 96));
  M2S_RC(grow(kBox, (size_t)nt * 16));
  M2S_RC(grow(kCnt, (size_t)nt * 4));
  // hipMemcpyDefault: the mesh may be host or device memory (unified addressing)
  if (nv) M2S_TRY(hipMemcpyAsync(b_[kInV].p, vertices, (size_t)nv * 12, hipMemcpyDefault, s_), "upload");
  if (nt) M2S_TRY(hipMemcpyAsync(b_[kInT].p, triangles, (size_t)nt * 12, hipMemcpyDefault, s_), "upload");
  const float* dV = at<float>(b_[kInV].p);
  const int32_t* dT = at<int32_t>(b_[kInT].p);
  if (nt)
    M2S_LAUNCH(k_tris, dim3((unsigned)blocks_for(nt, cu_)), dim3(256), 0, dV, nv, dT, nt, g,
               at<float4>(b_[kTri].p), at<uint4>(b_[kBox].p), at<uint32_t>(b_[kCnt].p), ctr);
  unsigned long long h[kCtrSlots];
  M2S_TRY(hipMemcpyAsync(h, ctr, sizeof(h), hipMemcpyDeviceToHost, s_), "counts");
  M2S_TRY(hipStreamSynchronize(s_), "counts");
  if (h[kCtrErr] & kErrIndex) return fail(DC_HIP_ERR_INVALID_ARG, "a triangle names a vertex outside [0, n_vertices)");
  if (h[kCtrErr] & kErrNonFinite) return fail(DC_HIP_ERR_NONFINITE, "a vertex coordinate is NaN or infinite");
  if (h[kCtrPairs] > kMaxPairs) return fail(DC_HIP_ERR_INVALID_ARG, "more than 2^32 - 1 triangle-bin pairs");
  const uint32_t pairs = (uint32_t)h[kCtrPairs];
  S.triangles = (int64_t)h[kCtrTris];
  S.bin_entries = pairs;
  S.bins = nbins;
  {
    // slack of the box test: a generous bound on the rounding of the FP32 distance
    uint32_t mbits = (uint32_t)h[kCtrMaxAbs];
    float m;
    std::memcpy(&m, &mbits, 4);
    for (int a = 0; a < 3; ++a) {
      m = std::fmax(m, std::fabs(origin[a]));
      m = std::fmax(m, std::fabs(origin[a] + (float)dims[a] * spacing[a]));
    }
    g.slack = 1e-5f * m + 1e-30f;
  }

  // per-axis tables, in the reference's FP32 operations (:247-251, :666-668)
  const int vres = voxelize_first ? std::max(32, std::min(2048, (int)voxel_res)) : 0;
  VoxGeom vg{};
  if (voxelize_first) {
    vg.vres = vres;
    vg.W = (vres + 63) / 64;
    vg.ox = g.ox;
    vg.oy = g.oy;
    vg.oz = g.oz;
    vg.sx = ((float)nx * g.vx) / (float)vres;
    vg.sy = ((float)ny * g.vy) / (float)vres;
    vg.sz = ((float)nz * g.vz) / (float)vres;
    const float reach = std::fmax(vg.sx, std::fmax(vg.sy, vg.sz)) * 0.8660254f;
    vg.reach_sq = reach * reach;
  }
  {
    const int n[3] = {nx, ny, nz}, gb[3] = {g.gx, g.gy, g.gz};
    const float o[3] = {g.ox, g.oy, g.oz}, v[3] = {g.vx, g.vy, g.vz}, bd[3] = {g.bdx, g.bdy, g.bdz};
    const float vs[3] = {vg.sx, vg.sy, vg.sz};
    axis_host_.assign((size_t)2 * (nx + ny + nz) + g.gx + g.gy + g.gz + 3, 0);
    size_t pos = 0;
    int32_t* bin_of[3];
    for (int a = 0; a < 3; ++a) {
      bin_of[a] = axis_host_.data() + pos;
      for (int i = 0; i < n[a]; ++i) {
        const float p = o[a] + (float)i * v[a];
        bin_of[a][i] = clamp_index((p - o[a]) / bd[a], gb[a]);
      }
      pos += n[a];
    }
    for (int a = 0; a < 3; ++a) {  // first sample of each bin (bin_of is non-decreasing)
      int32_t* st0 = axis_host_.data() + pos;
      int b = 0;
      for (int i = 0; i < n[a]; ++i)
        while (b <= bin_of[a][i]) st0[b++] = i;
      while (b <= gb[a]) st0[b++] = n[a];
      pos += gb[a] + 1;
    }
    for (int a = 0; a < 3; ++a) {
      int32_t* vx_of = axis_host_.data() + pos;
      if (voxelize_first)
        for (int i = 0; i < n[a]; ++i) {
          const float p = o[a] + (float)i * v[a];
          vx_of[i] = clamp_index((p - o[a]) / vs[a], vres);
        }
      pos += n[a];
    }
    M2S_RC(grow(kAxis, axis_host_.size() * 4));
    M2S_TRY(hipMemcpyAsync(b_[kAxis].p, axis_host_.data(), axis_host_.size() * 4,
                           hipMemcpyHostToDevice, s_),
            "axis tables");
  }
  Axis ax;
  {
    const int32_t* base = at<int32_t>(b_[kAxis].p);
    ax.bx = base;
    ax.by = ax.bx + nx;
    ax.bz = ax.by + ny;
    ax.sx = ax.bz + nz;
    ax.sy = ax.sx + g.gx + 1;
    ax.sz = ax.sy + g.gy + 1;
    ax.vx = ax.sz + g.gz + 1;
    ax.vy = ax.vx + nx;
    ax.vz = ax.vy + ny;
  }

  // bins: pairs, sorted by bin (stable), bin starts, evaluated bins
  M2S_RC(grow(kBinCnt, (size_t)(nbins + 1) * 4));
  M2S_RC(grow(kBinStart, (size_t)(nbins + 1) * 4));
  M2S_RC(grow(kBinAct, (size_t)nbins * 4));
  M2S_RC(grow(kActIdx, (size_t)nbins * 4));
  M2S_RC(grow(kActList, (size_t)nbins * 4));
  M2S_RC(grow(kTSum, (size_t)std::max<int64_t>(1, scan_tiles(std::max<int64_t>(nbins + 1, nt))) * 4));
  uint32_t* bincnt = at<uint32_t>(b_[kBinCnt].p);
  uint32_t* bstart = at<uint32_t>(b_[kBinStart].p);
  uint32_t* tsum = at<uint32_t>(b_[kTSum].p);
  uint32_t* scan_total = reinterpret_cast<uint32_t*>(ctr + kCtrActive);
  const uint32_t* tri_of = nullptr;
  M2S_LAUNCH(k_fill_u32, dim3((unsigned)blocks_for(nbins + 1, cu_)), dim3(256), 0, bincnt,
             nbins + 1, 0u);
  if (pairs > 0) {
    M2S_RC(grow(kOff, (size_t)nt * 4));
    M2S_RC(grow(kKeyA, (size_t)pairs * 4));
    M2S_RC(grow(kValA, (size_t)pairs * 4));
    M2S_RC(grow(kKeyB, (size_t)pairs * 4));
    M2S_RC(grow(kValB, (size_t)pairs * 4));
    uint32_t* off = at<uint32_t>(b_[kOff].p);
    M2S_TRY(exclusive_scan(at<uint32_t>(b_[kCnt].p), nt, off, tsum, scan_total, s_), "pair scan");
    uint32_t *ka = at<uint32_t>(b_[kKeyA].p), *va = at<uint32_t>(b_[kValA].p);
    uint32_t *kb = at<uint32_t>(b_[kKeyB].p), *vb = at<uint32_t>(b_[kValB].p);
    M2S_LAUNCH(k_pairs, dim3((unsigned)blocks_for(pairs, cu_)), dim3(256), 0, off, nt,
               at<uint4>(b_[kBox].p), pairs, g.gx, g.gy, ka, va);
    const uint32_t ntiles = (pairs + kSortTile - 1) / kSortTile;
    M2S_RC(grow(kHist, (size_t)ntiles * 256 * 4));
    M2S_RC(grow(kTSum, (size_t)std::max<int64_t>(
                           {int64_t(1), scan_tiles((int64_t)ntiles * 256), scan_tiles(nbins + 1),
                            scan_tiles(nt)}) * 4));
    tsum = at<uint32_t>(b_[kTSum].p);
    uint32_t* hist = at<uint32_t>(b_[kHist].p);
    for (int shift = 0; shift < 32 && ((uint64_t)(nbins - 1) >> shift) > 0; shift += 8) {
      M2S_LAUNCH(k_sort_hist, dim3(ntiles), dim3(256), 0, ka, pairs, shift, ntiles, hist);
      M2S_TRY(exclusive_scan(hist, (int64_t)ntiles * 256, hist, tsum, scan_total, s_), "sort scan");
      M2S_LAUNCH(k_sort_scatter, dim3(ntiles), dim3(256), 0, ka, va, pairs, shift, ntiles, hist,
                 kb, vb);
      std::swap(ka, kb);
      std::swap(va, vb);
    }
    tri_of = va;
    M2S_LAUNCH(k_bin_count, dim3((unsigned)blocks_for(pairs, cu_)), dim3(256), 0, ka, pairs,
               bincnt);
  }
  M2S_TRY(exclusive_scan(bincnt, nbins + 1, bstart, tsum, scan_total, s_), "bin scan");
  uint32_t* act = at<uint32_t>(b_[kBinAct].p);
  uint32_t* act_idx = at<uint32_t>(b_[kActIdx].p);
  uint32_t* act_list = at<uint32_t>(b_[kActList].p);
  M2S_LAUNCH(k_bin_active, dim3((unsigned)blocks_for(nbins, cu_)), dim3(256), 0, bincnt, g.gx,
             g.gy, g.gz, act);
  M2S_TRY(exclusive_scan(act, nbins, act_idx, tsum, scan_total, s_), "active scan");
  M2S_LAUNCH(k_compact, dim3((unsigned)blocks_for(nbins, cu_)), dim3(256), 0, act, act_idx, nbins,
             act_list);
  uint32_t n_act = 0;
  M2S_TRY(hipMemcpyAsync(&n_act, scan_total, 4, hipMemcpyDeviceToHost, s_), "active count");
  M2S_TRY(hipStreamSynchronize(s_), "active count");
  S.active_bins = n_act;

  hipEvent_t ev[2] = {nullptr, nullptr};
  M2S_TRY(hipEventCreate(&ev[0]), "event");
  M2S_TRY(hipEventCreate(&ev[1]), "event");
  struct EvGuard {
    hipEvent_t* e;
    ~EvGuard() {
      (void)hipEventDestroy(e[0]);
      (void)hipEventDestroy(e[1]);
    }
  } guard{ev};
  const dim3 sgrid((unsigned)((nx + 255) / 256), (unsigned)ny, (unsigned)nz);
  const dim3 rgrid((unsigned)((nx + 63) / 64), (unsigned)((ny + 3) / 4),
                   (unsigned)((nz + kRelaxRun - 1) / kRelaxRun));
  M2S_TRY(hipEventRecord(ev[0], s_), "event");
  M2S_LAUNCH(k_fill_far, sgrid, dim3(256), 0, g, ax, act, vals);
  if (n_act)
    M2S_LAUNCH(k_eval, dim3(n_act), dim3(256), 0, g, ax, act_list, bstart, tri_of,
               at<float4>(b_[kTri].p), vals);
  M2S_TRY(hipEventRecord(ev[1], s_), "event");

  int passes = 0;
  if (!voxelize_first) {
    const int W = (nx + 63) / 64;
    const size_t words = (size_t)W * ny * nz;
    M2S_RC(grow(kP, words * 8));
    M2S_RC(grow(kE, words * 8));
    M2S_RC(grow(kTmp, (size_t)N * 4));
    uint64_t* P = at<uint64_t>(b_[kP].p);
    uint64_t* E = at<uint64_t>(b_[kE].p);
    M2S_LAUNCH(k_sample_masks, dim3((unsigned)((W + 3) / 4), (unsigned)ny, (unsigned)nz),
               dim3(256), 0, vals, nx, ny, nz, W, P, E);
    M2S_RC(flood(P, E, nx, ny, nz, &passes));
    float* tmp = at<float>(b_[kTmp].p);
    const float keep = std::fmin(g.vx, std::fmin(g.vy, g.vz)) * 1.5f;
    M2S_LAUNCH(k_relax, rgrid, dim3(256), 0, vals, tmp, g, keep, (const uint64_t*)E, W);
    M2S_LAUNCH(k_relax, rgrid, dim3(256), 0, tmp, vals, g, keep, (const uint64_t*)nullptr, W);
    M2S_LAUNCH(k_relax, rgrid, dim3(256), 0, vals, tmp, g, keep, (const uint64_t*)nullptr, W);
    M2S_LAUNCH(k_relax, rgrid, dim3(256), 0, tmp, vals, g, keep, (const uint64_t*)nullptr, W);
  } else {
    const size_t words = (size_t)vg.W * vres * vres;
    M2S_RC(grow(kShell, words * 8));
    M2S_RC(grow(kP, words * 8));
    M2S_RC(grow(kE, words * 8));
    unsigned long long* shell = at<unsigned long long>(b_[kShell].p);
    uint64_t* P = at<uint64_t>(b_[kP].p);
    uint64_t* E = at<uint64_t>(b_[kE].p);
    M2S_LAUNCH(k_fill_u64, dim3((unsigned)blocks_for((int64_t)words, cu_)), dim3(256), 0, shell,
               (int64_t)words, 0ull);
    if (nt)
      M2S_LAUNCH(k_vox_raster, dim3((unsigned)blocks_for(nt, cu_)), dim3(256), 0, dV, dT, nt, vg,
                 shell);
    M2S_LAUNCH(k_vox_masks, dim3((unsigned)blocks_for((int64_t)words, cu_)), dim3(256), 0,
               (const uint64_t*)shell, vg, P, E);
    M2S_RC(flood(P, E, vres, vres, vres, &passes));
    M2S_LAUNCH(k_vox_apply, sgrid, dim3(256), 0, g, ax, (const uint64_t*)E, vg, vals);
  }
  if (!dev_out)
    M2S_TRY(hipMemcpyAsync(out, vals, (size_t)N * 4, hipMemcpyDefault, s_), "download");
  M2S_TRY(hipStreamSynchronize(s_), "synchronize");
  float eval_ms = 0.f;
  M2S_TRY(hipEventElapsedTime(&eval_ms, ev[0], ev[1]), "event");
  S.eval_ms = eval_ms;
  S.flood_passes = passes;
  S.voxel_res = vres;
  S.total_ms =
      std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (st) *st = S;
  return DC_HIP_OK;
}
