// dc_cells.hip — per-face and per-active-cell kernels of the dense dual-contouring path (gfx950).
//
//   k_pencil_words  per face word: index of its first face (pw), from the pencil bases
//   k_face_normals  per face: the 4-cell averaged hermite normal, once per grid edge
//   k_qef           per active cell: QEF vertex (FP64 plain, or FP32 with a list of cells to
//                   solve again in FP64, k_qef_fallback)
//   k_emit          per face: the oriented quad in the CPU backend's order
//   k_bwd_cell      adjoint per active cell, two phases (rows, then each cell's share of the
//                   averaged normals' adjoints)
//   k_bwd_dilate    candidate-sample plane of the adjoint
//   k_bwd_gather    dL/dphi per sample: fixed-order sum of the stencil contributions
//
// Faces are enumerated from the face planes one 64-bit word (64 k values of a pencil) per thread;
// a word's faces are consecutive in the CPU backend's face order (contouring.cpp:607-729).

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dc_cell.h"
#include "dc_common.h"
#include "dc_internal.h"
#include "dc_kernels.h"

namespace {

using namespace dccell;
using dccommon::find_bucket;

__device__ __forceinline__ void decode_cell(const DcGridDev& G, uint64_t entry, int& i, int& j,
                                            int& k) {
  const int64_t lw = (int64_t)(entry >> 6) - G.act_off;
  const int64_t row = lw / G.W;
  const int c = (int)(lw - row * G.W);
  i = 64 * c + (int)(entry & 63);
  j = (int)(row % G.cy);
  k = (int)(row / G.cy);
}

// ---------------------------------------------------------------------------------------------
// Face enumeration
// ---------------------------------------------------------------------------------------------
// First face index of every face word: pw = pbase(pencil) + faces in the pencil's earlier words.
__global__ void __launch_bounds__(256)
    k_pencil_words(const DcGridDev* __restrict__ grids, int n_grids,
                   const int64_t* __restrict__ pen_bounds, int64_t n_pen,
                   const uint64_t* __restrict__ ft, const uint32_t* __restrict__ pbase,
                   uint32_t* __restrict__ pw) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pen) return;
  const int g = find_bucket(pen_bounds, n_grids, p);
  const DcGridDev& G = grids[g];
  const int64_t local = p - G.pen_off;
  const int64_t plane = (int64_t)G.cy * G.cx;
  const int a = (int)(local / plane);
  const int64_t rem = local - (int64_t)a * plane;
  uint32_t pos = pbase[p];
  for (int kw = 0; kw < G.Wk; ++kw) {
    const int64_t idx = G.ft_off + ((int64_t)a * G.Wk + kw) * plane + rem;
    pw[idx] = pos;
    pos += (uint32_t)__popcll(ft[idx]);
  }
}

// fn(G, a, i, j, k, face) for every face of face word `idx` (one thread per word).
template <typename Fn>
__device__ __forceinline__ void word_faces(const DcGridDev* __restrict__ grids, int n_grids,
                                           const int64_t* __restrict__ ft_bounds, int64_t n_ft,
                                           const uint64_t* __restrict__ ft,
                                           const uint32_t* __restrict__ pw, Fn&& fn) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n_ft) return;
  uint64_t w = ft[idx];
  if (!w) return;
  const int g = find_bucket(ft_bounds, n_grids, idx);
  const DcGridDev& G = grids[g];
  const int64_t plane = (int64_t)G.cy * G.cx;
  const int64_t local = idx - G.ft_off;
  const int64_t aw = local / plane;  // a * Wk + kw
  const int64_t rem = local - aw * plane;
  const int a = (int)(aw / G.Wk);
  const int kw = (int)(aw - (int64_t)a * G.Wk);
  const int j = (int)(rem / G.cx);
  const int i = (int)(rem - (int64_t)j * G.cx);
  uint64_t face = pw[idx];
  while (w) {
    const int b = __ffsll((long long)w) - 1;
    w &= w - 1;
    fn(G, a, i, j, 64 * kw + b, face);
    ++face;
  }
}

// Averaged hermite normal of every face edge, once per face (plain: FP64 in the global frame;
// FP32 mode: FP32 in each cell's frame): normalized sum and the sum's length.
template <typename T>
__global__ void __launch_bounds__(256)
    k_face_normals(const DcGridDev* __restrict__ grids, int n_grids,
                   const int64_t* __restrict__ ft_bounds, int64_t n_ft,
                   const uint64_t* __restrict__ ft, const uint32_t* __restrict__ pw,
                   uint64_t capF, T* __restrict__ nrm) {
  word_faces(grids, n_grids, ft_bounds, n_ft, ft, pw,
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

// ---------------------------------------------------------------------------------------------
// Vertices
// ---------------------------------------------------------------------------------------------
// One thread per active cell (persistent grid-stride over the device-side count). In FP32 mode a
// cell that needs FP64 is appended to fblist (order irrelevant: results are per vertex).
template <int PREC>
__global__ void __launch_bounds__(256)
    k_qef(const DcGridDev* __restrict__ grids, int n_grids, const int64_t* __restrict__ act_bounds,
          const uint64_t* __restrict__ act, const uint64_t* __restrict__ alist,
          DcCounters* __restrict__ ctr, uint64_t cap, double* __restrict__ v64,
          float* __restrict__ v32, uint32_t* __restrict__ flags, uint32_t* __restrict__ fblist,
          const uint64_t* __restrict__ ft, const uint32_t* __restrict__ pw,
          const void* __restrict__ nrm, uint64_t capF) {
  const uint64_t n = min(ctr->n_active, cap);
  const uint64_t nf = min(ctr->n_faces, capF);
  const FaceNormals<double> fn64{ft, pw, reinterpret_cast<const double*>(nrm), nullptr, nf};
  const FaceNormals<float> fn32{ft, pw, nullptr, reinterpret_cast<const float*>(nrm), nf};
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t entry = alist[v];
    const int g = find_bucket(act_bounds, n_grids, (int64_t)(entry >> 6));
    const DcGridDev G = grids[g];
    int i, j, k;
    decode_cell(G, entry, i, j, k);
    if (PREC == DC_PREC_FP64) {
      double x[3];
      solve_cell_fp64(G, act, fn64, i, j, k, x);
      v64[3 * v + 0] = x[0];
      v64[3 * v + 1] = x[1];
      v64[3 * v + 2] = x[2];
      v32[3 * v + 0] = (float)x[0];
      v32[3 * v + 1] = (float)x[1];
      v32[3 * v + 2] = (float)x[2];
    } else {
      float xl[3];
      if (solve_cell_fp32(G, act, fn32, i, j, k, xl)) {
        v32[3 * v + 0] = (float)(coordd(G.ox, i, G.vx) + (double)xl[0]);
        v32[3 * v + 1] = (float)(coordd(G.oy, j, G.vy) + (double)xl[1]);
        v32[3 * v + 2] = (float)(coordd(G.oz, k, G.vz) + (double)xl[2]);
        flags[v] = 0u;
      } else {
        flags[v] = 1u;
        const uint64_t slot = atomicAdd((unsigned long long*)&ctr->fallback, 1ull);
        if (slot < cap) fblist[slot] = (uint32_t)v;
      }
    }
  }
}

// FP32 mode: the listed cells solved in FP64 (inline averaged normals, the plain arithmetic).
__global__ void __launch_bounds__(256)
    k_qef_fallback(const DcGridDev* __restrict__ grids, int n_grids,
                   const int64_t* __restrict__ act_bounds, const uint64_t* __restrict__ act,
                   const uint64_t* __restrict__ alist, const DcCounters* __restrict__ ctr,
                   uint64_t cap, const uint32_t* __restrict__ fblist, float* __restrict__ v32) {
  const uint64_t n = min(ctr->fallback, cap);
  const FaceNormals<double> inline64{nullptr, nullptr, nullptr, nullptr, 0ull};
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = fblist[t];
    const uint64_t entry = alist[v];
    const int g = find_bucket(act_bounds, n_grids, (int64_t)(entry >> 6));
    const DcGridDev G = grids[g];
    int i, j, k;
    decode_cell(G, entry, i, j, k);
    double x[3];
    solve_cell_fp64(G, act, inline64, i, j, k, x);
    v32[3 * (uint64_t)v + 0] = (float)x[0];
    v32[3 * (uint64_t)v + 1] = (float)x[1];
    v32[3 * (uint64_t)v + 2] = (float)x[2];
  }
}

// ---------------------------------------------------------------------------------------------
// Quads (contouring.cpp:607-729)
// ---------------------------------------------------------------------------------------------
// contouring.cpp:141-207 gradientAt: central differences with clamped neighbours, normalized.
__device__ __forceinline__ void grid_gradient(const DcGridDev& G, int i, int j, int k,
                                              double out[3]) {
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

// Vertex id (grid-local) of active cell (i,j,k): rank in (k,j,i) order.
__device__ __forceinline__ int32_t vertex_id(const DcGridDev& G, const uint64_t* __restrict__ act,
                                             const uint32_t* __restrict__ wprefix, uint32_t vbase,
                                             int i, int j, int k) {
  const int64_t w = G.act_off + ((int64_t)k * G.cy + j) * G.W + (i >> 6);
  const uint64_t below = (1ull << (i & 63)) - 1ull;
  return (int32_t)(wprefix[w] - vbase + (uint32_t)__popcll(act[w] & below));
}

__global__ void __launch_bounds__(256)
    k_emit(const DcGridDev* __restrict__ grids, int n_grids, const int64_t* __restrict__ ft_bounds,
           int64_t n_ft, const uint64_t* __restrict__ act, const uint32_t* __restrict__ wprefix,
           const uint64_t* __restrict__ ft, const uint32_t* __restrict__ pw,
           const double* __restrict__ v64, const float* __restrict__ v32, int precision,
           const DcCounters* __restrict__ ctr, uint64_t capA, uint64_t capF,
           int32_t* __restrict__ faces) {
  // an overflowed run is repeated by the host; never index past the buffers
  if (ctr->n_active > capA || ctr->n_faces > capF) return;
  word_faces(grids, n_grids, ft_bounds, n_ft, ft, pw,
             [&](const DcGridDev& G, int a, int i, int j, int k, uint64_t pos) {
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

// ---------------------------------------------------------------------------------------------
// Adjoint
// ---------------------------------------------------------------------------------------------
constexpr int kBwdThreads = 64;
constexpr int kStencil = 64;  // 4x4x4 samples around a cell: offsets -1..2 per axis

// Cells processed: every active cell (list == nullptr; with skip_flagged the FP32 kernel leaves
// the FP64 cells to the list launch) or the listed ones (the FP64 cells of FP32 mode).
template <int PHASE, typename T>
__global__ void __launch_bounds__(kBwdThreads)
    k_bwd_cell(const DcGridDev* __restrict__ grids, int n_grids,
               const int64_t* __restrict__ act_bounds, const uint64_t* __restrict__ act,
               const uint64_t* __restrict__ alist, const uint32_t* __restrict__ flags,
               int skip_flagged, const uint32_t* __restrict__ list,
               const DcCounters* __restrict__ ctr, uint64_t cap, const float* __restrict__ gv,
               const uint64_t* __restrict__ ft, const uint32_t* __restrict__ pw,
               const double* __restrict__ nrm64, const float* __restrict__ nrm32, uint64_t capF,
               float* __restrict__ ubar, float* __restrict__ contrib) {
  __shared__ float s_lc[kBwdThreads][kStencil + 1];
  float* lc = s_lc[threadIdx.x];
  const uint64_t n = list ? min(ctr->fallback, cap) : min(ctr->n_active, cap);
  const uint64_t nf = min(ctr->n_faces, capF);
  const FaceNormals<T> fn{ft, pw, nrm64, nrm32, nf};
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t v = list ? (uint64_t)list[t] : t;
    if (!list && skip_flagged && flags[v]) continue;
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
    if (PHASE == 1) {
      const float xb[3] = {gv[3 * v], gv[3 * v + 1], gv[3 * v + 2]};
      backward_rows<T>(G, act, fn, i, j, k, xb, lc, ubar);
    } else {
      backward_faces<T>(G, act, fn, i, j, k, ubar, lc);
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
  uint64_t am = 0, a0 = 0, ap = 0;  // OR of cell chunks c-1, c, c+1 over the 16 cell rows
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

// One wave per candidate-plane word (64 samples of a row): the 16 cell rows' words are loaded once
// per wave; each candidate lane sums its stencil contributions over cells in (k, j, i) order.
__global__ void __launch_bounds__(256)
    k_bwd_gather(const DcGridDev* __restrict__ grids, int n_grids,
                 const int64_t* __restrict__ dil_bounds, int64_t n_dil,
                 const uint64_t* __restrict__ act, const uint32_t* __restrict__ wprefix,
                 const uint64_t* __restrict__ dil, const float* __restrict__ contrib,
                 float* __restrict__ grad) {
  const int lane = threadIdx.x & 63;
  const int64_t wv = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (wv >= n_dil) return;  // uniform per wave
  const int g = find_bucket(dil_bounds, n_grids, wv);
  const DcGridDev G = grids[g];
  const int64_t local = wv - dil_bounds[g];
  const int64_t row = local / G.Ws;
  const int c = (int)(local - row * G.Ws);
  const int y = (int)(row % G.ny), z = (int)(row / G.ny);
  const int x = 64 * c + lane;
  const uint64_t d = dil[wv];
  float acc = 0.f;
  if (d) {
    const bool cand = ((d >> lane) & 1ull) && x < G.nx;
    for (int kz = max(z - 2, 0); kz <= min(z + 1, G.cz - 1); ++kz)
      for (int jy = max(y - 2, 0); jy <= min(y + 1, G.cy - 1); ++jy) {
        const int64_t rw = G.act_off + ((int64_t)kz * G.cy + jy) * G.W;
        const uint64_t wm = (c >= 1 && c - 1 < G.W) ? act[rw + c - 1] : 0ull;
        const uint64_t w0 = (c < G.W) ? act[rw + c] : 0ull;
        const uint64_t wp = (c + 1 < G.W) ? act[rw + c + 1] : 0ull;
        if (!(wm | w0 | wp)) continue;
        const uint32_t pm = (c >= 1 && c - 1 < G.W) ? wprefix[rw + c - 1] : 0u;
        const uint32_t p0 = (c < G.W) ? wprefix[rw + c] : 0u;
        const uint32_t pp = (c + 1 < G.W) ? wprefix[rw + c + 1] : 0u;
        if (!cand) continue;
        for (int dd = -2; dd <= 1; ++dd) {
          const int ix = x + dd;
          if (ix < 0 || ix >= G.cx) continue;
          const int sel = (ix >> 6) - c + 1;  // 0: chunk c-1, 1: c, 2: c+1
          const uint64_t word = sel == 0 ? wm : (sel == 1 ? w0 : wp);
          const uint32_t pre = sel == 0 ? pm : (sel == 1 ? p0 : pp);
          const uint64_t bit = 1ull << (ix & 63);
          if (!(word & bit)) continue;
          const uint64_t vid = (uint64_t)pre + (uint64_t)__popcll(word & (bit - 1ull));
          const int slot = (z - kz + 1) * 16 + (y - jy + 1) * 4 + (x - ix + 1);
          acc += contrib[vid * kStencil + slot];
        }
      }
  }
  if (x < G.nx) grad[G.samp_off + row * G.nx + x] = acc;
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------------------------
hipError_t dc_launch_face_normals(const DcFaceNormalArgs& a, hipStream_t s) {
  if (a.n_pen > 0)
    hipLaunchKernelGGL(k_pencil_words, dim3((unsigned)((a.n_pen + 255) / 256)), dim3(256), 0, s,
                       a.grids, a.n_grids, a.pen_bounds, a.n_pen, a.ft, a.pbase, a.pw);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || a.n_ft <= 0) return e;
  const dim3 grid((unsigned)((a.n_ft + 255) / 256)), block(256);
  if (a.precision == DC_PREC_FP64)
    hipLaunchKernelGGL(k_face_normals<double>, grid, block, 0, s, a.grids, a.n_grids,
                       a.ft_bounds, a.n_ft, a.ft, a.pw, a.capF, reinterpret_cast<double*>(a.nrm));
  else
    hipLaunchKernelGGL(k_face_normals<float>, grid, block, 0, s, a.grids, a.n_grids, a.ft_bounds,
                       a.n_ft, a.ft, a.pw, a.capF, reinterpret_cast<float*>(a.nrm));
  return hipGetLastError();
}

hipError_t dc_launch_qef(const DcQefArgs& a, hipStream_t s) {
  const dim3 grid((unsigned)a.blocks), block(256);
  if (a.precision == DC_PREC_FP64) {
    hipLaunchKernelGGL(k_qef<DC_PREC_FP64>, grid, block, 0, s, a.grids, a.n_grids, a.act_bounds,
                       a.act, a.alist, a.ctr_rw, a.capA, a.v64, a.v32, a.flags, a.fblist, a.ft,
                       a.pw, a.nrm, a.capF);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_qef<DC_PREC_FP32>, grid, block, 0, s, a.grids, a.n_grids, a.act_bounds,
                     a.act, a.alist, a.ctr_rw, a.capA, a.v64, a.v32, a.flags, a.fblist, a.ft, a.pw,
                     a.nrm, a.capF);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_qef_fallback, dim3((unsigned)std::max<int64_t>(1, a.blocks / 4)), block,
                     0, s, a.grids, a.n_grids, a.act_bounds, a.act, a.alist, a.ctr_rw, a.capA,
                     a.fblist, a.v32);
  return hipGetLastError();
}

hipError_t dc_launch_emit(const DcEmitArgs& a, hipStream_t s) {
  if (a.n_ft <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_emit, dim3((unsigned)((a.n_ft + 255) / 256)), dim3(256), 0, s, a.grids,
                     a.n_grids, a.ft_bounds, a.n_ft, a.act, a.wprefix, a.ft, a.pw, a.v64, a.v32,
                     a.precision, a.ctr, a.capA, a.capF, a.faces);
  return hipGetLastError();
}

hipError_t dc_launch_backward(const DcBwdArgs& a, hipStream_t s) {
  const double* n64 = a.precision == DC_PREC_FP64 ? reinterpret_cast<const double*>(a.nrm) : nullptr;
  const float* n32 = a.precision == DC_PREC_FP64 ? nullptr : reinterpret_cast<const float*>(a.nrm);
  const dim3 grid((unsigned)a.cell_blocks), block(kBwdThreads);
  const dim3 fgrid((unsigned)std::max<int64_t>(1, a.cell_blocks / 8));
  hipError_t e;
#define DC_BWD(PHASE)                                                                              \
  if (a.precision == DC_PREC_FP64) {                                                               \
    hipLaunchKernelGGL((k_bwd_cell<PHASE, double>), grid, block, 0, s, a.grids, a.n_grids,         \
                       a.act_bounds, a.act, a.alist, a.flags, 0, nullptr, a.ctr, a.capA,           \
                       a.grad_vertices, a.ft, a.pw, n64, n32, a.capF, a.ubar, a.contrib);          \
  } else {                                                                                         \
    hipLaunchKernelGGL((k_bwd_cell<PHASE, float>), grid, block, 0, s, a.grids, a.n_grids,          \
                       a.act_bounds, a.act, a.alist, a.flags, 1, nullptr, a.ctr, a.capA,           \
                       a.grad_vertices, a.ft, a.pw, n64, n32, a.capF, a.ubar, a.contrib);          \
    e = hipGetLastError();                                                                         \
    if (e != hipSuccess) return e;                                                                 \
    hipLaunchKernelGGL((k_bwd_cell<PHASE, double>), fgrid, block, 0, s, a.grids, a.n_grids,        \
                       a.act_bounds, a.act, a.alist, a.flags, 0, a.fblist, a.ctr, a.capA,          \
                       a.grad_vertices, a.ft, a.pw, n64, n32, a.capF, a.ubar, a.contrib);          \
  }                                                                                                \
  e = hipGetLastError();                                                                           \
  if (e != hipSuccess) return e;
  DC_BWD(1)
  DC_BWD(2)
#undef DC_BWD
  if (a.n_dil > 0) {
    hipLaunchKernelGGL(k_bwd_dilate, dim3((unsigned)((a.n_dil + 255) / 256)), dim3(256), 0, s,
                       a.grids, a.n_grids, a.dil_bounds, a.n_dil, a.act, a.dil);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_bwd_gather, dim3((unsigned)((a.n_dil + 3) / 4)), dim3(256), 0, s,
                       a.grids, a.n_grids, a.dil_bounds, a.n_dil, a.act, a.wprefix, a.dil,
                       a.contrib, a.grad_values);
  }
  return hipGetLastError();
}
