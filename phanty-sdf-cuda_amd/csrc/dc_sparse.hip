// dc_sparse.hip — sparse brick dual contouring and its multi-vertex form on CDNA4 (gfx950): the
// reference's CudaSparseDualContouringBackend / CudaSparseMvdcDualContouringBackend
// (src/cuda/dc_cuda_sparse.cu:1170-1485).
//
// The reference reads the grid once per brick (one thread scans (B+1)^3 samples), marks cells in
// a byte per cell, compacts bricks and cells with copy_if and keeps an int per cell for the
// vertex map. Here the grid is read once, coalesced, into three bit planes; bricks, cells and
// ranks are then bit work on those planes, and nothing is stored per cell of the grid:
//
//   planes   uint64 sample bit planes, 64 samples along x per word, word (k*ny + j)*Ws + i/64
//            (Ws = ceil(nx/64)): N (f < 0), P (f > 0), C (|f| <= 0.45 vx), in that order
//   rows     per brick row (bz, by) and sample word w: the OR of N (then of P) over the row's
//            sample rows j in [by*B, by*B + B], k in [bz*B, bz*B + B] (clipped), word
//            (bz*nby + by)*Ws + w; N rows first, then P rows
//   bricks   uint8 per brick (bx + nbx*(by + nby*bz)): 1 when its samples hold both signs
//   act      the active cells, one uint64 per 64 cells along x: word (k*cy + j)*W + i/64,
//            W = ceil((nx-1)/64), plus one zero word past the end
//   wprefix  uint32 per act word: active cells before it, so the rank of cell (i, j, k) is
//            wprefix[w] + popc(act[w] & below(i % 64)): rank order is linear cell order
//   alist    uint64 per active cell in rank order: act word * 64 + bit
//   one-vertex form: verts float[3] per active cell (vertex id = rank)
//   multi-vertex form: spos / snrm float[3] x 2 slots per active cell (a split cell's two
//            vertices and normals, slot 0 only for the others); vcnt uint32 per active cell and
//            vstart, its exclusive scan (the first vertex id); verts gathered from the slots
//   fcnt / fstart  uint32 per active cell: quads and their exclusive scan (first quad)
//   faces    int32[4] per quad
//
// FP32 arithmetic follows the reference's kernels operation by operation (built with
// -ffp-contract=off, IEEE division and square root), so every vertex equals the reference's.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>

#include "dc_common.h"
#include "dc_flagscan.h"
#include "dc_kernels.h"
#include "dc_sparse.h"

namespace {

using dccommon::wave_scan_incl;

// counters (uint32 words of kCtr); the act scan's total (uint64) sits at word kActive
enum {
  kNonfinite = 0,
  kBrickCount,
  kFallbacks,
  kClamps,
  kInvalid,
  kSplitRej,
  kSkipped,
  kVertTotal,
  kFaceTotal,
  kActive = 12,  // uint64 (words 12, 13)
  kNumCtr = 16
};

constexpr int kThreads = 256;
constexpr int kPlaneWords = 16;  // sample words per wave in the plane pass
constexpr int kCellRows = 8;    // cell rows (along y) per thread in the cell pass

struct SpGrid {
  const float* f;  // the samples (global memory)
  int nx, ny, nz;
  int cx, cy, cz, W;
  int64_t layer;   // nx * ny
  float vx, vy, vz, ox, oy, oz;
  float near;  // 0.45 vx
  __device__ __forceinline__ int64_t idx(int i, int j, int k) const {
    return (int64_t)k * layer + (int64_t)j * nx + i;
  }
  __device__ __forceinline__ float at(int64_t q) const { return dccommon::to_global(f)[q]; }
  __device__ __forceinline__ float at(int i, int j, int k) const { return at(idx(i, j, k)); }
};

struct F3 {
  float x, y, z;
};
__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 add(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ F3 sub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ F3 mul(float s, F3 a) { return f3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 cross(F3 a, F3 b) {
  return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ bool finite3(F3 a) {
  return isfinite(a.x) && isfinite(a.y) && isfinite(a.z);
}

// Corner q of a cell: offset (cox(q), coy(q), coz(q)); edge e joins corners ea(e) and eb(e)
// (dc_cuda_sparse.cu:477-486). Constant-folded in the unrolled loops below.
__device__ constexpr int cox(int q) { return (q == 1 || q == 2 || q == 5 || q == 6) ? 1 : 0; }
__device__ constexpr int coy(int q) { return (q == 2 || q == 3 || q == 6 || q == 7) ? 1 : 0; }
__device__ constexpr int coz(int q) { return q >= 4 ? 1 : 0; }
__device__ constexpr int ea(int e) { return e < 4 ? e : (e < 8 ? e : e - 8); }
__device__ constexpr int eb(int e) { return e < 4 ? (e + 1) % 4 : (e < 8 ? 4 + (e - 3) % 4 : e - 4); }

// dc_cuda_sparse.cu:151-176: central differences at a sample, one-sided at the border.
__device__ __forceinline__ F3 sample_gradient(const SpGrid& g, int i, int j, int k) {
  const int im = i > 0 ? i - 1 : 0, ip = i < g.nx - 1 ? i + 1 : g.nx - 1;
  const int jm = j > 0 ? j - 1 : 0, jp = j < g.ny - 1 ? j + 1 : g.ny - 1;
  const int km = k > 0 ? k - 1 : 0, kp = k < g.nz - 1 ? k + 1 : g.nz - 1;
  const int64_t b = g.idx(i, j, k);
  const float fim = g.at(b + (im - i)), fip = g.at(b + (ip - i));
  const float fjm = g.at(b + (int64_t)(jm - j) * g.nx), fjp = g.at(b + (int64_t)(jp - j) * g.nx);
  const float fkm = g.at(b + (km - k) * g.layer), fkp = g.at(b + (kp - k) * g.layer);
  F3 r = f3(0.f, 0.f, 0.f);
  const float dx = (float)(ip - im) * g.vx, dy = (float)(jp - jm) * g.vy,
              dz = (float)(kp - km) * g.vz;
  if (dx > 1e-12f) r.x = (fip - fim) / dx;
  if (dy > 1e-12f) r.y = (fjp - fjm) / dy;
  if (dz > 1e-12f) r.z = (fkp - fkm) / dz;
  return r;
}

__device__ __forceinline__ F3 normalized(F3 n) {
  const float l = sqrtf(dot(n, n));
  return l > 1e-6f ? mul(1.0f / l, n) : n;
}

// dc_cuda_sparse.cu:90-149: gradient of the cell's trilinear interpolant at p. f: the corner
// values; (x0, x1), (y0, y1), (z0, z1): the cell's corner coordinates per axis.
__device__ __forceinline__ F3 trilinear_gradient(const float f[8], float x0, float x1, float y0,
                                                 float y1, float z0, float z1, F3 p) {
  const float dx = x1 - x0, dy = y1 - y0, dz = z1 - z0;
  float u = 0.5f, v = 0.5f, w = 0.5f;
  if (fabsf(dx) > 1e-12f) u = (p.x - x0) / dx;
  if (fabsf(dy) > 1e-12f) v = (p.y - y0) / dy;
  if (fabsf(dz) > 1e-12f) w = (p.z - z0) / dz;
  u = fmaxf(0.0f, fminf(1.0f, u));
  v = fmaxf(0.0f, fminf(1.0f, v));
  w = fmaxf(0.0f, fminf(1.0f, w));
  const float ou = 1.0f - u, ov = 1.0f - v, ow = 1.0f - w;
  const float gu = (f[1] - f[0]) * ov * ow + (f[2] - f[3]) * v * ow + (f[5] - f[4]) * ov * w +
                   (f[6] - f[7]) * v * w;
  const float gv = (f[3] - f[0]) * ou * ow + (f[2] - f[1]) * u * ow + (f[7] - f[4]) * ou * w +
                   (f[6] - f[5]) * u * w;
  const float gw = (f[4] - f[0]) * ou * ov + (f[5] - f[1]) * u * ov + (f[7] - f[3]) * ou * v +
                   (f[6] - f[2]) * u * v;
  F3 r = f3(0.f, 0.f, 0.f);
  if (fabsf(dx) > 1e-12f) r.x = gu / dx;
  if (fabsf(dy) > 1e-12f) r.y = gv / dy;
  if (fabsf(dz) > 1e-12f) r.z = gw / dz;
  return r;
}

// The symmetric 3x3 held as its six distinct entries (the reference keeps all nine, written in
// mirrored pairs: dc_cuda_sparse.cu:226-236).
struct Sym3 {
  float a00, a01, a02, a11, a12, a22;
};

// dc_cuda_sparse.cu:178-269: at most 15 rotations on the largest upper off-diagonal entry (stop
// below 1e-7), then y = V S+ V^T d dropping singular values <= 0.01. Returns "every value dropped".
// The rotation plane (p, q) of a sweep is one of (0,1), (0,2), (1,2); its entries are picked and
// written back with selects (one code path for the three planes), with the reference's
// arithmetic on them (:214-242).
__device__ __forceinline__ bool jacobi_pinv(const Sym3& B, const float d[3], float y[3]) {
  float a00 = B.a00, a01 = B.a01, a02 = B.a02, a11 = B.a11, a12 = B.a12, a22 = B.a22;
  float V[3][3] = {{1.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, {0.f, 0.f, 1.f}};
  for (int it = 0; it < 15; ++it) {
    int sel = 0;
    float big = fabsf(a01);
    if (fabsf(a02) > big) {
      sel = 1;
      big = fabsf(a02);
    }
    if (fabsf(a12) > big) {
      sel = 2;
      big = fabsf(a12);
    }
    if (big < 1e-7f) break;
    const bool s0 = sel == 0, s1 = sel == 1;
    // (p, q, r) = (0, 1, 2), (0, 2, 1), (1, 2, 0)
    const float app = sel < 2 ? a00 : a11;
    const float aqq = s0 ? a11 : a22;
    const float apq = s0 ? a01 : (s1 ? a02 : a12);
    const float arp = s0 ? a02 : a01;
    const float arq = s0 ? a12 : (s1 ? a12 : a02);
    const float phi = 0.5f * (aqq - app) / apq;
    float t = 1.0f / (fabsf(phi) + sqrtf(phi * phi + 1.0f));
    if (phi < 0.0f) t = -t;
    const float c = 1.0f / sqrtf(t * t + 1.0f);
    const float s = t * c;
    const float rp = c * arp - s * arq;
    const float rq = s * arp + c * arq;
    const float pp = app - t * apq;
    const float qq = aqq + t * apq;
    if (s0) {
      a00 = pp; a11 = qq; a01 = 0.0f; a02 = rp; a12 = rq;
    } else if (s1) {
      a00 = pp; a22 = qq; a02 = 0.0f; a01 = rp; a12 = rq;
    } else {
      a11 = pp; a22 = qq; a12 = 0.0f; a01 = rp; a02 = rq;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const float vrp = sel < 2 ? V[r][0] : V[r][1];
      const float vrq = s0 ? V[r][1] : V[r][2];
      const float np = c * vrp - s * vrq;
      const float nq = s * vrp + c * vrq;
      V[r][0] = sel < 2 ? np : V[r][0];
      V[r][1] = s0 ? nq : (s1 ? V[r][1] : np);
      V[r][2] = s0 ? V[r][2] : nq;
    }
  }
  const float S[3] = {a00, a11, a22};
  float Si[3];
  int kept = 0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (S[q] > 0.01f) {
      Si[q] = 1.0f / S[q];
      ++kept;
    } else {
      Si[q] = 0.0f;
    }
  }
  float w[3];
  w[0] = V[0][0] * d[0] + V[1][0] * d[1] + V[2][0] * d[2];
  w[1] = V[0][1] * d[0] + V[1][1] * d[1] + V[2][1] * d[2];
  w[2] = V[0][2] * d[0] + V[1][2] * d[1] + V[2][2] * d[2];
  w[0] *= Si[0];
  w[1] *= Si[1];
  w[2] *= Si[2];
  y[0] = V[0][0] * w[0] + V[0][1] * w[1] + V[0][2] * w[2];
  y[1] = V[1][0] * w[0] + V[1][1] * w[1] + V[1][2] * w[2];
  y[2] = V[2][0] * w[0] + V[2][1] * w[1] + V[2][2] * w[2];
  return kept == 0;
}

// One cell's hermite entries in fixed slots: slot 2e (edge e's crossing entry, or a thin edge's
// first) and slot 2e+1 (a thin edge's second). Entry order is slot order; bit e of useA / useB
// says slot 2e / 2e+1 is an entry. The points of both slots of an edge are p[e]; slot 2e's normal
// is nA[e], slot 2e+1's is gc[eb(e)] (the normalized gradient of the edge's second corner).
struct Cell {
  float f[8];
  float x0, x1, y0, y1, z0, z1;
  F3 p[12], nA[12];
  F3 gc[8];
  uint32_t useA, useB;
  int m;
};

__device__ __forceinline__ F3 corner_pos(const Cell& c, int q) {
  return f3(cox(q) ? c.x1 : c.x0, coy(q) ? c.y1 : c.y0, coz(q) ? c.z1 : c.z0);
}

// dc_cuda_sparse.cu:488-563 / :682-758: the entries of cell (i, j, k).
template <int MODE>
__device__ __forceinline__ void cell_entries(const SpGrid& g, int i, int j, int k, Cell& c) {
  const int64_t b = g.idx(i, j, k);
#pragma unroll
  for (int q = 0; q < 8; ++q) c.f[q] = g.at(b + cox(q) + (int64_t)coy(q) * g.nx + coz(q) * g.layer);
  c.x0 = g.ox + (float)i * g.vx;
  c.x1 = g.ox + (float)(i + 1) * g.vx;
  c.y0 = g.oy + (float)j * g.vy;
  c.y1 = g.oy + (float)(j + 1) * g.vy;
  c.z0 = g.oz + (float)k * g.vz;
  c.z1 = g.oz + (float)(k + 1) * g.vz;
  bool cross[12], maybe[12];
  bool any_maybe = false;
#pragma unroll
  for (int e = 0; e < 12; ++e) {
    const float fa = c.f[ea(e)], fb = c.f[eb(e)];
    cross[e] = fa * fb <= 0.0f;
    maybe[e] = !cross[e] && !(fabsf(fa) > g.near || fabsf(fb) > g.near);
    any_maybe = any_maybe || maybe[e];
  }
  // raw corner gradients: for the thin-edge test (:310-319) and the edge-gradient normals
  F3 rg[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) rg[q] = f3(0.f, 0.f, 0.f);
  if (MODE == 1 || any_maybe) {
#pragma unroll
    for (int q = 0; q < 8; ++q) rg[q] = sample_gradient(g, i + cox(q), j + coy(q), k + coz(q));
  }
  bool thin[12];
  bool any_thin = false;
#pragma unroll
  for (int e = 0; e < 12; ++e) {
    thin[e] = maybe[e] && dot(rg[ea(e)], rg[eb(e)]) < -0.3f;
    any_thin = any_thin || thin[e];
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) c.gc[q] = f3(0.f, 0.f, 0.f);
  if (any_thin) {
#pragma unroll
    for (int q = 0; q < 8; ++q) c.gc[q] = normalized(rg[q]);
  }
  c.useA = c.useB = 0;
  c.m = 0;
#pragma unroll
  for (int e = 0; e < 12; ++e) {
    c.p[e] = f3(0.f, 0.f, 0.f);
    c.nA[e] = f3(0.f, 0.f, 0.f);
    if (cross[e] || thin[e]) {
      const float fa = c.f[ea(e)], fb = c.f[eb(e)];
      float t = cross[e] ? fa / (fa - fb + 1e-12f) : fabsf(fa) / (fabsf(fa) + fabsf(fb) + 1e-12f);
      t = fmaxf(0.0f, fminf(1.0f, t));
      const F3 pa = corner_pos(c, ea(e)), pb = corner_pos(c, eb(e));
      c.p[e] = add(mul(1.0f - t, pa), mul(t, pb));
      if (thin[e]) {
        c.nA[e] = c.gc[ea(e)];
        if (c.m < 11) {
          c.useA |= 1u << e;
          c.useB |= 1u << e;
          c.m += 2;
        }
      } else {
        F3 n;
        if (MODE == 1)
          n = add(mul(1.0f - t, rg[ea(e)]), mul(t, rg[eb(e)]));
        else
          n = trilinear_gradient(c.f, c.x0, c.x1, c.y0, c.y1, c.z0, c.z1, c.p[e]);
        c.nA[e] = normalized(n);
        if (c.m < 12) {
          c.useA |= 1u << e;
          c.m += 1;
        }
      }
    }
  }
}

struct Tally {
  uint32_t fallbacks = 0, clamps = 0, invalid = 0, split_rej = 0;
};

// dc_cuda_sparse.cu:380-443: the QEF of the entries in (tA, tB) about their mass point.
__device__ __forceinline__ F3 solve_cluster(const Cell& c, uint32_t tA, uint32_t tB,
                                            const SpGrid& g, Tally& tl) {
  const int cnt = __popc(tA) + __popc(tB);
  const F3 lo = f3(c.x0, c.y0, c.z0);
  const F3 centre = add(lo, mul(0.5f, f3(g.vx, g.vy, g.vz)));
  if (cnt == 0) return centre;
  F3 m = f3(0.f, 0.f, 0.f);
#pragma unroll
  for (int e = 0; e < 12; ++e) {
    if ((tA >> e) & 1) m = add(m, c.p[e]);
    if ((tB >> e) & 1) m = add(m, c.p[e]);
  }
  m = mul(1.0f / (float)cnt, m);
  Sym3 B = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  float d[3] = {0.f, 0.f, 0.f};
  auto acc = [&](F3 p, F3 n) {
    const float h = dot(n, sub(p, m));
    B.a00 += n.x * n.x;
    B.a01 += n.x * n.y;
    B.a02 += n.x * n.z;
    B.a11 += n.y * n.y;
    B.a12 += n.y * n.z;
    B.a22 += n.z * n.z;
    d[0] += n.x * h;
    d[1] += n.y * h;
    d[2] += n.z * h;
  };
#pragma unroll
  for (int e = 0; e < 12; ++e) {
    if ((tA >> e) & 1) acc(c.p[e], c.nA[e]);
    if ((tB >> e) & 1) acc(c.p[e], c.gc[eb(e)]);
  }
  float y[3];
  const bool singular = jacobi_pinv(B, d, y);
  F3 x = add(m, f3(y[0], y[1], y[2]));
  if (singular || !finite3(x)) {
    ++tl.fallbacks;
    x = m;
  }
  if (!finite3(x)) {
    ++tl.invalid;
    x = centre;
  }
  if (x.x < c.x0 || x.x > c.x1 || x.y < c.y0 || x.y > c.y1 || x.z < c.z0 || x.z > c.z1) {
    ++tl.clamps;
    x.x = fmaxf(c.x0, fminf(c.x1, x.x));
    x.y = fmaxf(c.y0, fminf(c.y1, x.y));
    x.z = fmaxf(c.z0, fminf(c.z1, x.z));
  }
  return x;
}

// Active-list entry -> cell (act words number fewer than 2^26: 32-bit arithmetic).
__device__ __forceinline__ void decode_cell(uint64_t entry, int W, int cy, int& i, int& j, int& k) {
  const uint32_t word = (uint32_t)(entry >> 6);
  const uint32_t row = word / (uint32_t)W;
  i = (int)(word - row * (uint32_t)W) * 64 + (int)(entry & 63);
  j = (int)(row % (uint32_t)cy);
  k = (int)(row / (uint32_t)cy);
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readlane((int)wave_scan_incl(v, threadIdx.x & 63), 63);
}

// Adds each of the block's N per-thread counts to its counter: one atomic per block and counter
// (none when the sum is zero). Every thread of the block calls it.
template <int N>
__device__ __forceinline__ void block_flush(const uint32_t (&v)[N], uint32_t* const (&dst)[N]) {
  __shared__ uint32_t s_part[N][kThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < N; ++q) {
    const uint32_t t = wave_sum(v[q]);
    if (lane == 0) s_part[q][w] = t;
  }
  __syncthreads();
  if (threadIdx.x < N) {
    uint32_t t = 0;
#pragma unroll
    for (int x = 0; x < kThreads / 64; ++x) t += s_part[threadIdx.x][x];
    if (t) atomicAdd(dst[threadIdx.x], t);
  }
}

// ---------------------------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------------------------

// The sample bit planes: one wave per kPlaneWords consecutive sample words (one sample per lane
// per word: 256-byte coalesced rows), all loads in flight before the ballots; non-finite samples
// flagged. The wave's position and row pointers are wave-uniform (scalar registers), so the vector
// work per word is the load, three compares and the finiteness test; the ballots reach the lanes
// that store them through LDS.
__global__ void __launch_bounds__(kThreads)
    k_sp_planes(SpGrid g, int Ws, int64_t n_words, uint64_t* __restrict__ N,
                uint64_t* __restrict__ P, uint64_t* __restrict__ C, uint32_t* __restrict__ ctr) {
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t q0 = ((int64_t)blockIdx.x * (kThreads / 64) + wv) * kPlaneWords;
  if (q0 >= n_words) return;  // uniform per wave
  const int nw = (int)min<int64_t>(kPlaneWords, n_words - q0);
  const int64_t row0 = q0 / Ws;
  int w = (int)(q0 - row0 * Ws);
  gfloat rp = dccommon::to_global(g.f) + row0 * g.nx;  // the current row (uniform)
  float v[kPlaneWords];
  uint64_t inm[kPlaneWords];  // lanes whose sample of word u lies in the row
#pragma unroll
  for (int u = 0; u < kPlaneWords; ++u) {
    inm[u] = 0;
    v[u] = 1.0f;
    if (u < nw) {  // uniform
      const int x = 64 * w + lane;
      inm[u] = __ballot(x < g.nx);
      v[u] = rp[min(x, g.nx - 1)];
    }
    if (++w == Ws) {
      w = 0;
      rp += g.nx;
    }
  }
  // the wave's plane words go through LDS to the lanes that store them: lane p*8 + u stores
  // word u of plane p (N, P, C)
  __shared__ uint64_t s_words[kThreads / 64][3 * kPlaneWords];
  uint64_t* sw = s_words[wv];
  uint64_t bad = 0;
#pragma unroll
  for (int u = 0; u < kPlaneWords; ++u) {
    const uint64_t bn = __ballot(v[u] < 0.0f) & inm[u];
    const uint64_t bp = __ballot(v[u] > 0.0f) & inm[u];
    const uint64_t bc = __ballot(fabsf(v[u]) <= g.near) & inm[u];
    bad |= __ballot(!isfinite(v[u])) & inm[u];
    if (lane == 0) {
      sw[u] = bn;
      sw[kPlaneWords + u] = bp;
      sw[2 * kPlaneWords + u] = bc;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (lane < 3 * kPlaneWords) {
    const int u = lane % kPlaneWords;
    if (u < nw) {
      uint64_t* dst = lane < kPlaneWords ? N : (lane < 2 * kPlaneWords ? P : C);
      dst[q0 + u] = sw[lane];
    }
  }
  if (bad != 0 && lane == 0) atomicOr(ctr + kNonfinite, 1u);
}

// Brick rows, in two steps. Thread (k, by, w) ORs word w of N and of P over the brick row's
// sample rows j in [by*B, by*B + B] of layer k; then thread (bz, by, w) ORs those over the
// layers k in [bz*B, bz*B + B].
__global__ void __launch_bounds__(kThreads)
    k_sp_rows_j(const uint64_t* __restrict__ N, const uint64_t* __restrict__ P, int ny, int Ws,
                int B, int nby, int64_t n, uint64_t* __restrict__ JN, uint64_t* __restrict__ JP) {
  const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (t >= n) return;
  const int w = (int)(t % Ws);
  const int64_t r = t / Ws;
  const int by = (int)(r % nby);
  const int64_t k = r / nby;
  const int j0 = by * B, j1 = min(j0 + B + 1, ny);
  uint64_t on = 0, op = 0;
  for (int j = j0; j < j1; ++j) {
    const int64_t q = (k * ny + j) * Ws + w;
    on |= N[q];
    op |= P[q];
  }
  JN[t] = on;
  JP[t] = op;
}

__global__ void __launch_bounds__(kThreads)
    k_sp_rows_k(const uint64_t* __restrict__ JN, const uint64_t* __restrict__ JP, int nz, int Ws,
                int B, int nby, int64_t n, uint64_t* __restrict__ RN, uint64_t* __restrict__ RP) {
  const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (t >= n) return;
  const int w = (int)(t % Ws);
  const int64_t r = t / Ws;
  const int by = (int)(r % nby), bz = (int)(r / nby);
  const int k0 = bz * B, k1 = min(k0 + B + 1, nz);
  uint64_t on = 0, op = 0;
  for (int k = k0; k < k1; ++k) {
    const int64_t q = ((int64_t)k * nby + by) * Ws + w;
    on |= JN[q];
    op |= JP[q];
  }
  RN[t] = on;
  RP[t] = op;
}

__device__ __forceinline__ uint64_t bit_range(int lo, int hi) {  // bits [lo, hi), 0 <= lo < hi <= 64
  const uint64_t top = hi >= 64 ? ~0ull : (1ull << hi) - 1ull;
  return top & ~((1ull << lo) - 1ull);
}

// dc_cuda_sparse.cu:271-308: brick b is active when its samples [bx*B, bx*B + B] (clipped) of
// its brick row hold a negative and a positive value.
__global__ void __launch_bounds__(kThreads)
    k_sp_brick_flags(const uint64_t* __restrict__ RN, const uint64_t* __restrict__ RP, int nx,
                     int Ws, int B, int nbx, int64_t n_bricks, uint8_t* __restrict__ flag,
                     uint32_t* __restrict__ ctr) {
  uint32_t cnt = 0;
  for (int64_t b = (int64_t)blockIdx.x * kThreads + threadIdx.x; b < n_bricks;
       b += (int64_t)gridDim.x * kThreads) {
    const int bx = (int)(b % nbx);
    const int64_t r = b / nbx;
    const int x0 = bx * B, x1 = min(x0 + B + 1, nx);
    uint64_t an = 0, ap = 0;
    for (int w = x0 >> 6; w <= (x1 - 1) >> 6; ++w) {
      const uint64_t m = bit_range(max(x0, 64 * w) - 64 * w, min(x1, 64 * w + 64) - 64 * w);
      an |= RN[r * Ws + w] & m;
      ap |= RP[r * Ws + w] & m;
    }
    const bool act = an != 0 && ap != 0;
    flag[b] = act ? 1 : 0;
    cnt += act;
  }
  const uint32_t v[1] = {cnt};
  uint32_t* const dst[1] = {ctr + kBrickCount};
  block_flush<1>(v, dst);
}

// Active cells of the brick row (by, bz) in act word w: the cells of active bricks.
__device__ __forceinline__ uint64_t brick_mask(const uint8_t* __restrict__ flag, int w, int cx,
                                               int B, int nbx, int64_t row_base) {
  const int c0 = 64 * w, c1 = min(64 * w + 64, cx);
  uint64_t m = 0;
  for (int bx = c0 / B; bx <= (c1 - 1) / B; ++bx)
    if (flag[row_base + bx]) m |= bit_range(max(bx * B, c0) - c0, min(bx * B + B, c1) - c0);
  return m;
}

// dc_cuda_sparse.cu:321-378: a cell of an active brick is active when its corners hold both
// signs or one within 0.45 vx of zero. Thread (k, row chunk, w) walks kCellRows cell rows along
// y, carrying the shared sample rows; a cell's corner bits are its sample bit and the next one.
__global__ void __launch_bounds__(kThreads)
    k_sp_cells(const uint64_t* __restrict__ N, const uint64_t* __restrict__ P,
               const uint64_t* __restrict__ C, const uint8_t* __restrict__ flag, int ny, int Ws,
               int W, int cx, int cy, int B, int nbx, int nby, int chunks, int64_t n,
               uint64_t* __restrict__ act) {
  const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (t >= n) return;
  const int w = (int)(t % W);
  const int64_t r = t / W;
  const int jc = (int)(r % chunks), k = (int)(r / chunks);
  const int j0 = jc * kCellRows, j1 = min(j0 + kCellRows, cy);
  const bool nxt = w + 1 < Ws;
  const int64_t l0 = (int64_t)k * ny, l1 = (int64_t)(k + 1) * ny;
  auto row = [&](const uint64_t* pl, int j, uint64_t& a, uint64_t& b) {
    a = pl[(l0 + j) * Ws + w] | pl[(l1 + j) * Ws + w];
    b = nxt ? (pl[(l0 + j) * Ws + w + 1] | pl[(l1 + j) * Ws + w + 1]) : 0ull;
  };
  uint64_t n0, n1, p0, p1, c0, c1;
  row(N, j0, n0, n1);
  row(P, j0, p0, p1);
  row(C, j0, c0, c1);
  const uint64_t valid = cx - 64 * w >= 64 ? ~0ull : (1ull << (cx - 64 * w)) - 1ull;
  const int bz = k / B;
  int by_cur = -1;
  uint64_t bm = 0;
  for (int j = j0; j < j1; ++j) {
    uint64_t n0b, n1b, p0b, p1b, c0b, c1b;
    row(N, j + 1, n0b, n1b);
    row(P, j + 1, p0b, p1b);
    row(C, j + 1, c0b, c1b);
    const uint64_t rn = n0 | n0b, rn1 = n1 | n1b;
    const uint64_t rp = p0 | p0b, rp1 = p1 | p1b;
    const uint64_t rc = c0 | c0b, rc1 = c1 | c1b;
    const uint64_t cn = rn | (rn >> 1) | (rn1 << 63);
    const uint64_t cp = rp | (rp >> 1) | (rp1 << 63);
    const uint64_t cc = rc | (rc >> 1) | (rc1 << 63);
    const int by = j / B;
    if (by != by_cur) {
      bm = brick_mask(flag, w, cx, B, nbx, ((int64_t)bz * nby + by) * nbx);
      by_cur = by;
    }
    act[((int64_t)k * cy + j) * W + w] = ((cn & cp) | cc) & valid & bm;
    n0 = n0b;
    n1 = n1b;
    p0 = p0b;
    p1 = p1b;
    c0 = c0b;
    c1 = c1b;
  }
}

// dc_cuda_sparse.cu:445-827: per active cell (rank r) its hermite entries, in the multi-vertex
// form the split (:565-620) and one solve per cluster with the clusters' mean normals
// (:786-826), else one solve.
template <int MODE, bool MV>
__device__ __forceinline__ void sp_vertices(SpGrid g, const uint64_t* __restrict__ alist,
                                            int64_t na, float* __restrict__ verts,
                                            float* __restrict__ spos, float* __restrict__ snrm,
                                            uint32_t* __restrict__ vcnt,
                                            uint32_t* __restrict__ ctr) {
  const int64_t r = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  Tally tl;
  const bool live = r < na;
  Cell c;
  if (live) {
    int i, j, k;
    decode_cell(alist[r], g.W, g.cy, i, j, k);
    cell_entries<MODE>(g, i, j, k, c);
  } else {
    c.useA = c.useB = 0;
    c.m = 0;
  }
  const uint32_t all = c.useA, allB = c.useB;
  if (!MV) {
    if (live) {
      const F3 x = solve_cluster(c, all, allB, g, tl);
      verts[3 * r] = x.x;
      verts[3 * r + 1] = x.y;
      verts[3 * r + 2] = x.z;
    }
  } else {
    // the split decision (:565-620): the entry normals in entry order go to LDS (one column per
    // thread), so the pair search and the sides run over this cell's m entries only
    __shared__ float s_n[3 * 12 * kThreads];
    auto put = [&](int q, F3 n) {
      s_n[(3 * q) * kThreads + threadIdx.x] = n.x;
      s_n[(3 * q + 1) * kThreads + threadIdx.x] = n.y;
      s_n[(3 * q + 2) * kThreads + threadIdx.x] = n.z;
    };
    auto get = [&](int q) {
      return f3(s_n[(3 * q) * kThreads + threadIdx.x], s_n[(3 * q + 1) * kThreads + threadIdx.x],
                s_n[(3 * q + 2) * kThreads + threadIdx.x]);
    };
    int nv = 1;
    uint32_t t0A = all, t0B = allB;
    F3 vA = f3(0.f, 0.f, 0.f), vB = f3(0.f, 0.f, 0.f);
    if (live && c.m >= 2) {
      int q = 0;
#pragma unroll
      for (int e = 0; e < 12; ++e) {
        if ((all >> e) & 1) put(q++, c.nA[e]);
        if ((allB >> e) & 1) put(q++, c.gc[eb(e)]);
      }
      float lo = 1.0f;
      int ia = -1, ib = -1;
      for (int x = 0; x < c.m; ++x) {
        const F3 nx_ = get(x);
        for (int y = x + 1; y < c.m; ++y) {
          const float t = dot(nx_, get(y));
          if (t < lo) {
            lo = t;
            ia = x;
            ib = y;
          }
        }
      }
      if (ia >= 0 && ib >= 0 && !(lo >= -0.3f)) {
        vA = get(ia);
        vB = get(ib);
        uint32_t side = 0;  // bit q: entry q is nearer nA
        int n0 = 0;
        for (int x = 0; x < c.m; ++x) {
          const F3 n = get(x);
          if (dot(n, vA) >= dot(n, vB)) {
            side |= 1u << x;
            ++n0;
          }
        }
        if (n0 < 2 || c.m - n0 < 2) {
          ++tl.split_rej;
        } else {
          nv = 2;
          // entry bits back to slot masks (entries are in slot order)
          uint32_t sA = 0, sB = 0;
          int x = 0;
#pragma unroll
          for (int e = 0; e < 12; ++e) {
            if ((all >> e) & 1) {
              sA |= ((side >> x) & 1u) << e;
              ++x;
            }
            if ((allB >> e) & 1) {
              sB |= ((side >> x) & 1u) << e;
              ++x;
            }
          }
          t0A = sA;
          t0B = sB;
        }
      }
    }
    if (live) {
      if (nv == 1) {
        const F3 x = solve_cluster(c, all, allB, g, tl);
        spos[6 * r] = x.x;
        spos[6 * r + 1] = x.y;
        spos[6 * r + 2] = x.z;
      } else {
        const uint32_t t1A = all & ~t0A, t1B = allB & ~t0B;
        const F3 x0 = solve_cluster(c, t0A, t0B, g, tl);
        const F3 x1 = solve_cluster(c, t1A, t1B, g, tl);
        F3 s0 = f3(0.f, 0.f, 0.f), s1 = f3(0.f, 0.f, 0.f);
#pragma unroll
        for (int e = 0; e < 12; ++e) {
          if ((t0A >> e) & 1) s0 = add(s0, c.nA[e]);
          if ((t1A >> e) & 1) s1 = add(s1, c.nA[e]);
          if ((t0B >> e) & 1) s0 = add(s0, c.gc[eb(e)]);
          if ((t1B >> e) & 1) s1 = add(s1, c.gc[eb(e)]);
        }
        const float l0 = sqrtf(dot(s0, s0)), l1 = sqrtf(dot(s1, s1));
        const F3 n0 = l0 > 1e-6f ? mul(1.0f / l0, s0) : vA;
        const F3 n1 = l1 > 1e-6f ? mul(1.0f / l1, s1) : vB;
        spos[6 * r] = x0.x;
        spos[6 * r + 1] = x0.y;
        spos[6 * r + 2] = x0.z;
        spos[6 * r + 3] = x1.x;
        spos[6 * r + 4] = x1.y;
        spos[6 * r + 5] = x1.z;
        snrm[6 * r] = n0.x;
        snrm[6 * r + 1] = n0.y;
        snrm[6 * r + 2] = n0.z;
        snrm[6 * r + 3] = n1.x;
        snrm[6 * r + 4] = n1.y;
        snrm[6 * r + 5] = n1.z;
      }
      vcnt[r] = (uint32_t)nv;
    }
  }
  // one- and two-vertex cells follow from the counts (two = vertices - cells)
  const uint32_t v[4] = {tl.fallbacks, tl.clamps, tl.invalid, tl.split_rej};
  uint32_t* const dst[4] = {ctr + kFallbacks, ctr + kClamps, ctr + kInvalid, ctr + kSplitRej};
  block_flush<4>(v, dst);
}

// The one-vertex form held to 128 registers (four waves per SIMD); the multi-vertex form keeps
// its split state in registers beside the LDS normals.
template <int MODE>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4)))
    k_sp_vertices1(SpGrid g, const uint64_t* __restrict__ alist, int64_t na,
                   float* __restrict__ verts, uint32_t* __restrict__ ctr) {
  sp_vertices<MODE, false>(g, alist, na, verts, nullptr, nullptr, nullptr, ctr);
}

template <int MODE>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3)))
    k_sp_vertices2(SpGrid g, const uint64_t* __restrict__ alist, int64_t na,
                   float* __restrict__ spos, float* __restrict__ snrm, uint32_t* __restrict__ vcnt,
                   uint32_t* __restrict__ ctr) {
  sp_vertices<MODE, true>(g, alist, na, nullptr, spos, snrm, vcnt, ctr);
}

// Multi-vertex form: the vertices in id order from the slots.
__global__ void __launch_bounds__(kThreads)
    k_sp_place(const float* __restrict__ spos, const uint32_t* __restrict__ vcnt,
               const uint32_t* __restrict__ vstart, int64_t na, float* __restrict__ verts) {
  const int64_t r = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (r >= na) return;
  const uint32_t n = vcnt[r];
  const int64_t v = vstart[r];
  for (uint32_t h = 0; h < n; ++h) {
    verts[3 * (v + h)] = spos[6 * r + 3 * h];
    verts[3 * (v + h) + 1] = spos[6 * r + 3 * h + 1];
    verts[3 * (v + h) + 2] = spos[6 * r + 3 * h + 2];
  }
}

struct FaceArgs {
  const uint64_t* act;
  const uint32_t* wpre;
  const uint64_t* alist;
  int64_t na;
  const float* verts;  // one-vertex form: by vertex id
  const float* spos;   // multi-vertex form: slots
  const float* snrm;
  const uint32_t* vcnt;
  const uint32_t* vstart;
  uint32_t* ctr;
};

// Rank of cell (i, j, k) in the active list, or -1 when it is not active.
__device__ __forceinline__ int cell_rank(const SpGrid& g, const FaceArgs& a, int i, int j, int k) {
  const uint32_t w = ((uint32_t)k * (uint32_t)g.cy + (uint32_t)j) * (uint32_t)g.W + (uint32_t)(i >> 6);
  const uint64_t word = a.act[w];
  const int bit = i & 63;
  if (!((word >> bit) & 1)) return -1;
  return (int)(a.wpre[w] + (uint32_t)__popcll(word & ((1ull << bit) - 1ull)));
}

// dc_cuda_sparse.cu:829-851: which of a split cell's two vertices a face whose summed edge
// gradients are gn takes (the one whose normal is nearer gn), or -1 when neither normal is within
// 120 degrees of gn.
__device__ __forceinline__ int pick_vertex(const FaceArgs& a, int rank, F3 gn) {
  const float* nr = a.snrm + 6 * (int64_t)rank;
  const float d0 = dot(gn, f3(nr[0], nr[1], nr[2]));
  const float d1 = dot(gn, f3(nr[3], nr[4], nr[5]));
  if (fmaxf(d0, d1) < -0.5f) return -1;
  return d0 >= d1 ? 0 : 1;
}

constexpr int kFacesPerCell = 6;                       // 3 edges, 2 quads for a thin edge
constexpr int kStagePerWave = 64 * kFacesPerCell;      // quads a wave can emit

// dc_cuda_sparse.cu:853-1102: the quads of active cell r: its x, y and z edge from its first
// corner, a strict sign change giving a quad and a thin edge the quad and its reverse, each of
// the four cells around the edge giving a vertex (:871-874, :894-897, :917-920), oriented along
// the summed end gradients. *skipped counts the edges that lacked a vertex.
template <bool MV>
__device__ __forceinline__ int cell_faces(const SpGrid& g, const FaceArgs& a, int64_t r,
                                          int4 (&fq)[kFacesPerCell], uint32_t* skipped) {
  int i, j, k;
  decode_cell(a.alist[r], g.W, g.cy, i, j, k);
  int nf = 0;
  const int64_t b = g.idx(i, j, k);
  const float f0 = g.at(b);
#pragma unroll
  for (int axis = 0; axis < 3; ++axis) {
    const bool has = axis == 0 ? (j >= 1 && k >= 1) : axis == 1 ? (i >= 1 && k >= 1)
                                                                 : (i >= 1 && j >= 1);
    if (!has) continue;
    const int i1 = i + (axis == 0), j1 = j + (axis == 1), k1 = k + (axis == 2);
    const float f1 = g.at(b + (axis == 0 ? 1 : axis == 1 ? (int64_t)g.nx : g.layer));
    const bool crossing = f0 * f1 < 0.0f;
    const bool maybe = !(f0 * f1 <= 0.0f) && !(fabsf(f0) > g.near || fabsf(f1) > g.near);
    if (!crossing && !maybe) continue;
    const F3 g0 = sample_gradient(g, i, j, k), g1 = sample_gradient(g, i1, j1, k1);
    const bool thin = !crossing && dot(g0, g1) < -0.3f;
    if (!crossing && !thin) continue;
    const F3 gn = add(g0, g1);
    int ci[4], cj[4], ck[4];
    if (axis == 0) {
      ci[0] = i; cj[0] = j - 1; ck[0] = k - 1;
      ci[1] = i; cj[1] = j;     ck[1] = k - 1;
      ci[2] = i; cj[2] = j;     ck[2] = k;
      ci[3] = i; cj[3] = j - 1; ck[3] = k;
    } else if (axis == 1) {
      ci[0] = i - 1; cj[0] = j; ck[0] = k - 1;
      ci[1] = i;     cj[1] = j; ck[1] = k - 1;
      ci[2] = i;     cj[2] = j; ck[2] = k;
      ci[3] = i - 1; cj[3] = j; ck[3] = k;
    } else {
      ci[0] = i - 1; cj[0] = j - 1; ck[0] = k;
      ci[1] = i;     cj[1] = j - 1; ck[1] = k;
      ci[2] = i;     cj[2] = j;     ck[2] = k;
      ci[3] = i - 1; cj[3] = j;     ck[3] = k;
    }
    int rk[4], v[4], h[4] = {0, 0, 0, 0};
    bool ok = true;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      rk[q] = cell_rank(g, a, ci[q], cj[q], ck[q]);
      ok = ok && rk[q] >= 0;
    }
    if (ok) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (MV && a.vcnt[rk[q]] == 2) {
          h[q] = pick_vertex(a, rk[q], gn);
          ok = ok && h[q] >= 0;
        }
        v[q] = MV ? (int)a.vstart[rk[q]] + h[q] : rk[q];
      }
    }
    if (!ok) {
      ++*skipped;
      continue;
    }
    F3 p[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float* src = MV ? a.spos + 6 * (int64_t)rk[q] + 3 * h[q] : a.verts + 3 * (int64_t)rk[q];
      p[q] = f3(src[0], src[1], src[2]);
    }
    F3 qn = cross(sub(p[1], p[0]), sub(p[3], p[0]));
    if (dot(qn, qn) < 1e-12f) qn = cross(sub(p[2], p[0]), sub(p[3], p[0]));
    if (dot(gn, gn) > 0.0f && dot(qn, qn) > 0.0f && dot(qn, gn) < 0.0f) {
      const int t = v[1];
      v[1] = v[3];
      v[3] = t;
    }
    // static slots: the quads of edge `axis` go to slots 2 axis and 2 axis + 1, then move down
    fq[2 * axis] = make_int4(v[0], v[1], v[2], v[3]);
    fq[2 * axis + 1] = make_int4(v[0], v[3], v[2], v[1]);
    nf |= (thin ? 3 : 1) << (2 * axis);
  }
  return nf;  // bits 2a, 2a + 1: slots of edge a in use
}

// The quads of every active cell in one pass: each wave writes its cells' quads, in cell order,
// to its own staging run (kStagePerWave quads) and its count to wcount; a scan of the counts and
// k_sp_face_copy then place the runs. The reference counts, scans and emits in three passes
// (:1359-1392); the order of the quads is the same.
template <bool MV>
__global__ void __launch_bounds__(kThreads) k_sp_faces(SpGrid g, FaceArgs a, int4* __restrict__ stage,
                                                       uint32_t* __restrict__ wcount) {
  const int64_t r = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const int lane = threadIdx.x & 63;
  uint32_t skipped = 0;
  int4 fq[kFacesPerCell];
  int used = 0;
  if (r < a.na) used = cell_faces<MV>(g, a, r, fq, &skipped);
  const uint32_t n = (uint32_t)__popc(used);
  const uint32_t incl = wave_scan_incl(n, lane);
  const int64_t wave = r >> 6;
  int4* out = stage + wave * kStagePerWave + (incl - n);
  int q = 0;
#pragma unroll
  for (int s = 0; s < kFacesPerCell; ++s)
    if ((used >> s) & 1) out[q++] = fq[s];
  if (lane == 63 && wave * 64 < a.na) wcount[wave] = incl;
  // the reference counts an edge without its four vertices in its emission pass, which skips
  // cells without quads (:953, :1008-1010)
  if (MV) {
    const uint32_t v[1] = {n ? skipped : 0u};
    uint32_t* const dst[1] = {a.ctr + kSkipped};
    block_flush<1>(v, dst);
  }
}

// One block of 64 threads per staging run: the run's quads to their place in the face order.
__global__ void __launch_bounds__(64)
    k_sp_face_copy(const int4* __restrict__ stage, const uint32_t* __restrict__ wcount,
                   const uint32_t* __restrict__ wbase, int4* __restrict__ faces) {
  const int64_t w = blockIdx.x;
  const uint32_t n = wcount[w], base = wbase[w];
  for (uint32_t q = threadIdx.x; q < n; q += 64) faces[base + q] = stage[w * kStagePerWave + q];
}

int64_t blocks_for(int64_t n, int64_t per_block) { return (n + per_block - 1) / per_block; }

}  // namespace

// ---------------------------------------------------------------------------------------------
// Host orchestration
// ---------------------------------------------------------------------------------------------
DcSparse::~DcSparse() {
  for (Buf& b : b_)
    if (b.p) (void)hipFree(b.p);
  if (h_ctr_) (void)hipHostFree(h_ctr_);
  for (hipEvent_t& e : ev_)
    if (e) (void)hipEventDestroy(e);
}

int DcSparse::hip_error(hipError_t e, const char* what) {
  if (err_) *err_ = std::string("sparse extraction: ") + what + ": " + hipGetErrorString(e);
  (void)hipGetLastError();
  return e == hipErrorOutOfMemory ? DC_HIP_ERR_OUT_OF_MEMORY : DC_HIP_ERR_HIP;
}

#define SP_TRY(expr, what)                            \
  do {                                                \
    const hipError_t e_ = (expr);                     \
    if (e_ != hipSuccess) return hip_error(e_, what); \
  } while (0)
#define SP_RC(expr)                   \
  do {                                \
    const int rc_ = (expr);           \
    if (rc_ != DC_HIP_OK) return rc_; \
  } while (0)
#define SP_LAUNCH(kernel, blocks, ...)                                                     \
  do {                                                                                     \
    if ((blocks) > 0) {                                                                    \
      hipLaunchKernelGGL(kernel, dim3((unsigned)(blocks)), dim3(kThreads), 0, s_, __VA_ARGS__); \
      SP_TRY(hipGetLastError(), #kernel);                                                  \
    }                                                                                      \
  } while (0)

int DcSparse::grow(int which, size_t bytes) {
  Buf& b = b_[which];
  if (bytes == 0) bytes = 256;
  if (b.cap >= bytes) return DC_HIP_OK;
  if (b.p) {
    SP_TRY(hipStreamSynchronize(s_), "synchronize");  // earlier work may still use it
    SP_TRY(hipFree(b.p), "free");
    b.p = nullptr;
    b.cap = 0;
  }
  const size_t want = bytes + bytes / 8;
  SP_TRY(hipMalloc(&b.p, want), "allocate");
  b.cap = want;
  return DC_HIP_OK;
}

template <typename T>
static T* at(void* p) {
  return static_cast<T*>(p);
}

int DcSparse::run(const dc_hip_grid& grid, bool on_device, int brick_size, int normal_mode,
                  bool multi_vertex, int cu_count, hipStream_t s, dc_hip_sparse_stats* st,
                  std::string* err) {
  const auto t_start = std::chrono::steady_clock::now();
  s_ = s;
  cu_ = std::max(1, cu_count);
  err_ = err;
  have_ = false;
  nv_ = nf_ = 0;
  auto fail = [&](int code, const char* msg) {
    if (err_) *err_ = msg;
    return code;
  };
  if (!grid.values) return fail(DC_HIP_ERR_INVALID_ARG, "grid values pointer is NULL");
  if (brick_size < 1) return fail(DC_HIP_ERR_INVALID_ARG, "brick_size must be >= 1");
  if (normal_mode < 0 || normal_mode > 2)
    return fail(DC_HIP_ERR_INVALID_ARG, "normal_mode must be 0, 1 or 2");
  const int nx = grid.nx, ny = grid.ny, nz = grid.nz;
  if (nx < 2 || ny < 2 || nz < 2)
    return fail(DC_HIP_ERR_INVALID_GRID, "grid dimensions must be >= 2");
  if (nx > 65536 || ny > 65536 || nz > 65536)
    return fail(DC_HIP_ERR_INVALID_GRID, "grid dimension above 65536");
  const int B = brick_size;
  const int cx = nx - 1, cy = ny - 1, cz = nz - 1;
  const int Ws = (nx + 63) / 64, W = (cx + 63) / 64;
  const int nbx = (cx + B - 1) / B, nby = (cy + B - 1) / B, nbz = (cz + B - 1) / B;
  const int64_t n_samples = (int64_t)nx * ny * nz;
  const int64_t n_sw = (int64_t)nz * ny * Ws;  // sample words per plane
  const int64_t n_words = (int64_t)cz * cy * W;
  const int64_t n_rows = (int64_t)nbz * nby * Ws;
  const int64_t n_bricks = (int64_t)nbx * nby * nbz;
  if ((uint64_t)n_words * 64ull >= (1ull << 32))
    return fail(DC_HIP_ERR_INVALID_GRID, "grid has too many cells for 32-bit ranks");

  if (!h_ctr_) SP_TRY(hipHostMalloc(reinterpret_cast<void**>(&h_ctr_), kNumCtr * 4), "pinned counters");
  for (hipEvent_t& e : ev_)
    if (!e) SP_TRY(hipEventCreate(&e), "event");
  auto mark = [&](int q) { return hipEventRecord(ev_[q], s_); };

  SP_TRY(mark(0), "event");
  const float* vals = grid.values;
  if (!on_device) {
    SP_RC(grow(kUp, (size_t)n_samples * 4));
    SP_TRY(hipMemcpyAsync(b_[kUp].p, grid.values, (size_t)n_samples * 4, hipMemcpyHostToDevice, s_),
           "upload");
    vals = at<float>(b_[kUp].p);
  }
  SP_TRY(mark(1), "event");

  SpGrid g{};
  g.f = vals;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.cx = cx;
  g.cy = cy;
  g.cz = cz;
  g.W = W;
  g.layer = (int64_t)nx * ny;
  g.vx = grid.vx;
  g.vy = grid.vy;
  g.vz = grid.vz;
  g.ox = grid.ox;
  g.oy = grid.oy;
  g.oz = grid.oz;
  g.near = 0.45f * grid.vx;

  SP_RC(grow(kPlanes, (size_t)n_sw * 3 * 8));
  const int64_t n_jrows = (int64_t)nz * nby * Ws;  // first step of the brick rows
  SP_RC(grow(kRows, (size_t)(n_rows + n_jrows) * 2 * 8));
  SP_RC(grow(kBricks, (size_t)n_bricks));
  SP_RC(grow(kAct, (size_t)(n_words + 1) * 8));
  SP_RC(grow(kWpre, (size_t)(n_words + 1) * 4));
  const int64_t tiles = dc_scan_tiles(n_words);
  SP_RC(grow(kScanSum, (size_t)tiles * 4));
  SP_RC(grow(kScanBase, (size_t)tiles * 8));
  SP_RC(grow(kCtr, kNumCtr * 4));
  SP_RC(grow(kAlist, 8));
  uint64_t* planes = at<uint64_t>(b_[kPlanes].p);
  uint64_t *N = planes, *P = planes + n_sw, *C = planes + 2 * n_sw;
  uint64_t* rows = at<uint64_t>(b_[kRows].p);
  uint8_t* bricks = at<uint8_t>(b_[kBricks].p);
  uint64_t* act = at<uint64_t>(b_[kAct].p);
  uint32_t* wpre = at<uint32_t>(b_[kWpre].p);
  uint32_t* ctr = at<uint32_t>(b_[kCtr].p);
  uint64_t* n_active_dev = reinterpret_cast<uint64_t*>(ctr + kActive);

  SP_TRY(hipMemsetAsync(ctr, 0, kNumCtr * 4, s_), "clear counters");
  SP_TRY(hipMemsetAsync(act + n_words, 0, 8, s_), "clear act tail");
  const int64_t plane_waves = (n_sw + kPlaneWords - 1) / kPlaneWords;
  SP_LAUNCH(k_sp_planes, blocks_for(plane_waves, kThreads / 64), g, Ws, n_sw, N, P, C, ctr);
  SP_TRY(mark(2), "event");
  uint64_t* jrows = rows + 2 * n_rows;
  SP_LAUNCH(k_sp_rows_j, blocks_for(n_jrows, kThreads), N, P, ny, Ws, B, nby, n_jrows, jrows,
            jrows + n_jrows);
  SP_LAUNCH(k_sp_rows_k, blocks_for(n_rows, kThreads), jrows, jrows + n_jrows, nz, Ws, B, nby,
            n_rows, rows, rows + n_rows);
  SP_LAUNCH(k_sp_brick_flags, std::min<int64_t>(blocks_for(n_bricks, kThreads), 512), rows,
            rows + n_rows, nx, Ws, B, nbx, n_bricks, bricks, ctr);
  SP_TRY(mark(3), "event");
  const int chunks = (cy + kCellRows - 1) / kCellRows;
  const int64_t n_cell_threads = (int64_t)cz * chunks * W;
  SP_LAUNCH(k_sp_cells, blocks_for(n_cell_threads, kThreads), N, P, C, bricks, ny, Ws, W, cx, cy,
            B, nbx, nby, chunks, n_cell_threads, act);
  SP_TRY(mark(4), "event");
  auto scan_act = [&]() -> int {
    SP_TRY(dc_launch_scan_act(act, n_words, at<uint32_t>(b_[kScanSum].p),
                              at<uint64_t>(b_[kScanBase].p), wpre, at<uint64_t>(b_[kAlist].p),
                              (uint64_t)(b_[kAlist].cap / 8), n_active_dev, s_),
           "active scan");
    return DC_HIP_OK;
  };
  SP_RC(scan_act());
  SP_TRY(mark(5), "event");
  SP_TRY(hipMemcpyAsync(h_ctr_, ctr, kNumCtr * 4, hipMemcpyDeviceToHost, s_), "read counts");
  SP_TRY(hipStreamSynchronize(s_), "active cells");
  if (h_ctr_[kNonfinite]) return fail(DC_HIP_ERR_NONFINITE, "grid contains NaN or Inf samples");
  uint64_t na_u;
  memcpy(&na_u, h_ctr_ + kActive, 8);
  const int64_t na = (int64_t)na_u;
  const int64_t active_bricks = h_ctr_[kBrickCount];
  if (na >= (1ll << 30)) return fail(DC_HIP_ERR_INVALID_GRID, "mesh exceeds 32-bit vertex ids");
  if ((uint64_t)na > b_[kAlist].cap / 8) {  // the list did not fit: size it and scan again
    SP_RC(grow(kAlist, (size_t)na * 8));
    SP_RC(scan_act());
  }
  const uint64_t* alist = at<uint64_t>(b_[kAlist].p);

  int64_t nv = na, nf = 0;
  float t_q = 0.f, t_fc = 0.f, t_fs = 0.f, t_ff = 0.f;
  if (na > 0) {
    SP_RC(grow(kVerts, (size_t)(multi_vertex ? 2 : 1) * na * 12));
    const int64_t n_waves = (na + 63) / 64;
    SP_RC(grow(kFcnt, (size_t)n_waves * 4));
    SP_RC(grow(kFstart, (size_t)n_waves * 4));
    SP_RC(grow(kStage, (size_t)n_waves * kStagePerWave * 16));
    SP_RC(grow(kTsum, (size_t)std::max<int64_t>(1, dcscan::scan_tiles(na)) * 4));
    if (multi_vertex) {
      SP_RC(grow(kSlotPos, (size_t)na * 24));
      SP_RC(grow(kSlotNrm, (size_t)na * 24));
      SP_RC(grow(kVcnt, (size_t)na * 4));
      SP_RC(grow(kVstart, (size_t)na * 4));
    }
    float* verts = at<float>(b_[kVerts].p);
    float* spos = multi_vertex ? at<float>(b_[kSlotPos].p) : nullptr;
    float* snrm = multi_vertex ? at<float>(b_[kSlotNrm].p) : nullptr;
    uint32_t* vcnt = multi_vertex ? at<uint32_t>(b_[kVcnt].p) : nullptr;
    uint32_t* vstart = multi_vertex ? at<uint32_t>(b_[kVstart].p) : nullptr;
    uint32_t* tsum = at<uint32_t>(b_[kTsum].p);
    const int64_t blk = blocks_for(na, kThreads);
    SP_TRY(mark(6), "event");
    if (multi_vertex) {
      if (normal_mode == 1)
        SP_LAUNCH(k_sp_vertices2<1>, blk, g, alist, na, spos, snrm, vcnt, ctr);
      else
        SP_LAUNCH(k_sp_vertices2<0>, blk, g, alist, na, spos, snrm, vcnt, ctr);
      SP_TRY(dcscan::exclusive_scan(vcnt, na, vstart, tsum, ctr + kVertTotal, s_), "vertex scan");
      SP_LAUNCH(k_sp_place, blk, spos, vcnt, vstart, na, verts);
    } else {
      if (normal_mode == 1)
        SP_LAUNCH(k_sp_vertices1<1>, blk, g, alist, na, verts, ctr);
      else
        SP_LAUNCH(k_sp_vertices1<0>, blk, g, alist, na, verts, ctr);
    }
    SP_TRY(mark(7), "event");
    FaceArgs fa;
    fa.act = act;
    fa.wpre = wpre;
    fa.alist = alist;
    fa.na = na;
    fa.verts = verts;
    fa.spos = spos;
    fa.snrm = snrm;
    fa.vcnt = vcnt;
    fa.vstart = vstart;
    fa.ctr = ctr;
    int4* stage = at<int4>(b_[kStage].p);
    uint32_t* wcount = at<uint32_t>(b_[kFcnt].p);
    uint32_t* wbase = at<uint32_t>(b_[kFstart].p);
    if (multi_vertex)
      SP_LAUNCH(k_sp_faces<true>, blk, g, fa, stage, wcount);
    else
      SP_LAUNCH(k_sp_faces<false>, blk, g, fa, stage, wcount);
    SP_TRY(mark(8), "event");
    SP_TRY(dcscan::exclusive_scan(wcount, n_waves, wbase, tsum, ctr + kFaceTotal, s_), "face scan");
    SP_TRY(mark(9), "event");
    SP_TRY(hipMemcpyAsync(h_ctr_, ctr, kNumCtr * 4, hipMemcpyDeviceToHost, s_), "read counts");
    SP_TRY(hipStreamSynchronize(s_), "face count");
    nf = h_ctr_[kFaceTotal];
    if (multi_vertex) nv = h_ctr_[kVertTotal];
    if (nf >= (1ll << 31) / 4) return fail(DC_HIP_ERR_INVALID_GRID, "mesh exceeds 32-bit face ids");
    SP_RC(grow(kFaces, (size_t)std::max<int64_t>(nf, 1) * 16));
    SP_TRY(mark(10), "event");
    if (nf > 0) {
      hipLaunchKernelGGL(k_sp_face_copy, dim3((unsigned)n_waves), dim3(64), 0, s_, stage, wcount,
                         wbase, at<int4>(b_[kFaces].p));
      SP_TRY(hipGetLastError(), "k_sp_face_copy");
    }
    SP_TRY(mark(11), "event");
    SP_TRY(hipMemcpyAsync(h_ctr_, ctr, kNumCtr * 4, hipMemcpyDeviceToHost, s_), "read counters");
    SP_TRY(hipStreamSynchronize(s_), "faces");
    auto ms = [&](int a, int b) {
      float v = 0.f;
      (void)hipEventElapsedTime(&v, ev_[a], ev_[b]);
      return v;
    };
    t_q = ms(6, 7);
    t_fc = ms(7, 8);
    t_fs = ms(8, 9);
    t_ff = ms(10, 11);
  }
  have_ = true;
  nv_ = nv;
  nf_ = nf;
  if (st) {
    auto ms = [&](int a, int b) {
      float v = 0.f;
      (void)hipEventElapsedTime(&v, ev_[a], ev_[b]);
      return v;
    };
    *st = dc_hip_sparse_stats{};
    st->nx = nx;
    st->ny = ny;
    st->nz = nz;
    st->brick_size = B;
    st->total_cells = (int64_t)cx * cy * cz;
    st->active_cells = na;
    st->total_bricks = n_bricks;
    st->active_bricks = active_bricks;
    st->vertex_count = nv;
    st->face_count = nf;
    st->qef_fallback_count = h_ctr_[kFallbacks];
    st->clamp_count = h_ctr_[kClamps];
    st->invalid_count = h_ctr_[kInvalid];
    // the multi-vertex form: every active cell has one or two vertices
    st->two_vertex_cells = multi_vertex ? nv - na : 0;
    st->ambiguous_cells = st->two_vertex_cells;
    st->multi_vertex_cells = st->two_vertex_cells;
    st->one_vertex_cells = multi_vertex ? na - st->two_vertex_cells : 0;
    st->split_rejection_count = multi_vertex ? h_ctr_[kSplitRej] : 0;
    st->bad_qef_count = 0;
    st->faces_skipped_due_to_missing_cluster = multi_vertex ? h_ctr_[kSkipped] : 0;
    st->upload_ms = on_device ? 0.f : ms(0, 1);
    st->sample_pass_ms = ms(1, 2);
    st->active_brick_marking_ms = ms(1, 3);
    st->active_brick_compaction_ms = 0.f;
    st->active_cell_marking_ms = ms(3, 4);
    st->active_cell_compaction_ms = ms(4, 5);
    st->marking_ms = st->active_brick_marking_ms + st->active_cell_marking_ms;
    st->compaction_ms = st->active_cell_compaction_ms;
    st->qef_ms = t_q;
    st->face_count_ms = t_fc;
    st->face_prefix_sum_ms = t_fs;
    st->face_fill_ms = t_ff;
    st->face_emission_ms = t_fc + t_fs + t_ff;
    st->download_ms = 0.f;
    st->total_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() -
                                                            t_start)
                       .count();
  }
  return DC_HIP_OK;
}
