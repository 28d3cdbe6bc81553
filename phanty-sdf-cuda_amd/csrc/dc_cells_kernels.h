// dc_cells_kernels.h — per-face and per-active-cell kernels of the dense dual-contouring path
// (gfx950), included by two translation units: dc_cells.hip (built with -ffp-contract=off: the
// plain FP64 mode, whose vertices and orientations must match the CPU backend bit for bit) and
// dc_cells_fast.hip (contraction allowed: the FP32 mode, its FP64 fallback cells and the
// adjoint, all held to tolerances). Each launches only its own instantiations.
//
//   k_pencil_words  per pencil: index of the first face of each face word (pw) and the face list
//   k_face_normals  per face: the 4-cell averaged hermite normal, once per grid edge
//   k_qef           per active cell: QEF vertex (FP64 plain, or FP32 with a list of cells to
//                   solve again in FP64, k_qef_fallback)
//   k_emit          per face: the oriented quad in the CPU backend's order
//   k_bwd_cell      adjoint per active cell, two phases (rows, then each cell's share of the
//                   averaged normals' adjoints), into the cell's own 8 corners
//   k_bwd_gather    dL/dphi per sample: fixed-order sum over the (up to) 8 cells containing it
//
// Faces are numbered in the CPU backend's face order (contouring.cpp:607-729) by the pencil scan;
// the per-face kernels run one thread per face-list entry.

#pragma once
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
// Per pencil (a,i,j), taken in (a, j, i) order so ft and pw accesses coalesce: the first face
// index of each of its face words (pw).
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
  const int64_t rem = local - (int64_t)a * plane;  // j * cx + i
  const int j = (int)(rem / G.cx), i = (int)(rem - (int64_t)j * G.cx);
  uint32_t pos = pbase[G.pen_off + ((int64_t)a * G.cx + i) * G.cy + j];  // face order
  for (int kw = 0; kw < G.Wk; ++kw) {
    const int64_t idx = G.ft_off + ((int64_t)a * G.Wk + kw) * plane + rem;
    pw[idx] = pos;
    pos += (uint32_t)__popcll(ft[idx]);
  }
}

__device__ __forceinline__ void decode_face(const DcGridDev& G, uint64_t entry, int& a, int& i,
                                            int& j, int& k) {
  const int64_t plane = (int64_t)G.cy * G.cx;
  const int64_t local = (int64_t)(entry >> 6) - G.ft_off;
  const int64_t aw = local / plane;  // a * Wk + kw
  const int64_t rem = local - aw * plane;
  a = (int)(aw / G.Wk);
  const int kw = (int)(aw - (int64_t)a * G.Wk);
  j = (int)(rem / G.cx);
  i = (int)(rem - (int64_t)j * G.cx);
  k = 64 * kw + (int)(entry & 63);
}

// fn(G, a, i, j, k, face id) for every face, one thread per face-list entry (grid-stride over
// the count).
template <typename Fn>
__device__ __forceinline__ void each_face(const DcGridDev* __restrict__ grids, int n_grids,
                                          const int64_t* __restrict__ ft_bounds,
                                          const uint64_t* __restrict__ flist,
                                          const uint64_t* __restrict__ ft,
                                          const uint32_t* __restrict__ pw, uint64_t n, Fn&& fn) {
  for (uint64_t f = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; f < n;
       f += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t entry = flist[f];
    const int64_t word = (int64_t)(entry >> 6);
    const uint64_t below = (1ull << (entry & 63)) - 1ull;
    const uint64_t face = (uint64_t)pw[word] + (uint64_t)__popcll(ft[word] & below);
    const int g = find_bucket(ft_bounds, n_grids, word);
    const DcGridDev G = grids[g];
    int a, i, j, k;
    decode_face(G, entry, a, i, j, k);
    fn(G, a, i, j, k, face);
  }
}

// Averaged hermite normal of every face edge, once per face (plain: FP64 in the global frame;
// FP32 mode: FP32 in each cell's frame): normalized sum and the sum's length.
template <typename T>
__global__ void __launch_bounds__(256)
    k_face_normals(const DcGridDev* __restrict__ grids, int n_grids,
                   const int64_t* __restrict__ ft_bounds, const uint64_t* __restrict__ flist,
                   const uint64_t* __restrict__ ft, const uint32_t* __restrict__ pw,
                   const DcCounters* __restrict__ ctr, uint64_t capF, T* __restrict__ nrm) {
  each_face(grids, n_grids, ft_bounds, flist, ft, pw, min(ctr->n_faces, capF),
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
           const uint64_t* __restrict__ flist, const uint64_t* __restrict__ ft,
           const uint32_t* __restrict__ pw, const uint64_t* __restrict__ act,
           const uint32_t* __restrict__ wprefix, const double* __restrict__ v64,
           const float* __restrict__ v32, int precision, const DcCounters* __restrict__ ctr,
           uint64_t capA, uint64_t capF, int32_t* __restrict__ faces) {
  // an overflowed run is repeated by the host; never index past the buffers
  if (ctr->n_active > capA || ctr->n_faces > capF) return;
  each_face(grids, n_grids, ft_bounds, flist, ft, pw, ctr->n_faces,
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
// Cells processed: every active cell (list == nullptr; with skip_flagged the FP32 kernel leaves
// the FP64 cells to the list launch) or the listed ones (the FP64 cells of FP32 mode). The
// cell's 8 corner adjoints go to contrib[8v..8v+7] (phase 1 writes, phase 2 adds).
// FP32 instances fit 4 waves per SIMD without spilling; FP64 ones keep the default bound.
template <int PHASE, typename T>
__global__ void __launch_bounds__(256, sizeof(T) == 4 ? 4 : 1)
    k_bwd_cell(const DcGridDev* __restrict__ grids, int n_grids,
               const int64_t* __restrict__ act_bounds, const uint64_t* __restrict__ act,
               const uint32_t* __restrict__ wprefix, const uint64_t* __restrict__ alist,
               const uint32_t* __restrict__ flags, int skip_flagged,
               const uint32_t* __restrict__ list, const DcCounters* __restrict__ ctr, uint64_t cap,
               const float* __restrict__ gv, const uint64_t* __restrict__ ft,
               const uint32_t* __restrict__ pw, const double* __restrict__ nrm64,
               const float* __restrict__ nrm32, uint64_t capF, float* __restrict__ ubar,
               float* __restrict__ ubar_e, float* __restrict__ contrib) {
  const uint64_t n = list ? min(ctr->fallback, cap) : min(ctr->n_active, cap);
  const uint64_t nf = min(ctr->n_faces, capF);
  const FaceNormals<T> fn{ft, pw, nrm64, nrm32, nf};
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t v = list ? (uint64_t)list[t] : t;
    if (!list && skip_flagged && flags[v]) continue;
    const uint64_t entry = alist[v];
    const int g = find_bucket(act_bounds, n_grids, (int64_t)(entry >> 6));
    const DcGridDev G = grids[g];
    int i, j, k;
    decode_cell(G, entry, i, j, k);
    T Sb[8];
    float4* out = reinterpret_cast<float4*>(contrib + 8 * v);
    if (PHASE == 1) {
      const float xb[3] = {gv[3 * v], gv[3 * v + 1], gv[3 * v + 2]};
      backward_rows<T>(G, act, fn, i, j, k, v, xb, ubar, ubar_e, Sb);
      out[0] = make_float4((float)Sb[0], (float)Sb[1], (float)Sb[2], (float)Sb[3]);
      out[1] = make_float4((float)Sb[4], (float)Sb[5], (float)Sb[6], (float)Sb[7]);
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) Sb[q] = T(0);
      backward_faces<T>(G, act, wprefix, fn, i, j, k, ubar, ubar_e, Sb);
      const float4 o0 = out[0], o1 = out[1];
      out[0] = make_float4(o0.x + (float)Sb[0], o0.y + (float)Sb[1], o0.z + (float)Sb[2],
                           o0.w + (float)Sb[3]);
      out[1] = make_float4(o1.x + (float)Sb[4], o1.y + (float)Sb[5], o1.z + (float)Sb[6],
                           o1.w + (float)Sb[7]);
    }
  }
}

// Corner index (contouring.cpp:232-241 numbering) at offset (dx,dy,dz) in {0,1}^3.
__device__ __forceinline__ int corner_at(int dx, int dy, int dz) {
  return 4 * dz + (dy ? (dx ? 2 : 3) : dx);
}

// dL/dphi, one wave per sample row (y, z): sample x sums the corner contributions of the active
// cells (x-1..x, y-1..y, z-1..z) in (k, j, i) order, and every sample is written.
// The row's 4 cell rows ((k, j) order; q bit 0: j = y, bit 1: k = z) are fetched with one vector
// load of their act words and vertex prefixes (lane = 16q + chunk, for up to 16 chunks per
// step). A row with no active cell is zero-filled. Otherwise, per chunk of 64 samples and cell
// row, the lane of cell x loads the cell's two corner values on this sample row with one float2
// load (dx = 0 and 1 are adjacent in the corner numbering); the dx = 1 value moves one lane up (to
// sample x+1) and lane 63's into the next chunk. Chunk words with no active cell are skipped on a
// uniform branch, and a step's contribution loads are all issued before any sum.
constexpr int kGatherChunks = 8;   // chunks whose loads are in flight together
constexpr int kGatherStep = 16;    // chunks per act-word load

__global__ void __launch_bounds__(256)
    k_bwd_gather(const DcGridDev* __restrict__ grids, int n_grids,
                 const int64_t* __restrict__ sr_bounds, int64_t n_sr,
                 const uint64_t* __restrict__ act, const uint32_t* __restrict__ wprefix,
                 const float* __restrict__ contrib, float* __restrict__ grad) {
  const int lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * (blockDim.x >> 6) +
                    __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (r >= n_sr) return;  // uniform per wave
  const int g = find_bucket(sr_bounds, n_grids, r);
  const DcGridDev G = grids[g];
  const int64_t local = r - G.sr_off;
  const int y = (int)(local % G.ny), z = (int)(local / G.ny);
  float* out = grad + G.samp_off + local * G.nx;
  const int lq = lane >> 4;
  const int jy = y - 1 + (lq & 1), kz = z - 1 + (lq >> 1);
  const bool row_ok = jy >= 0 && jy < G.cy && kz >= 0 && kz < G.cz;
  // clamped row so that every lane's (unconditional) prefix load stays inside the grid
  const int64_t rw = G.act_off + ((int64_t)min(max(kz, 0), G.cz - 1) * G.cy +
                                  min(max(jy, 0), G.cy - 1)) * G.W;
  const uint64_t below = (1ull << lane) - 1ull;
  float carry[4] = {0.f, 0.f, 0.f, 0.f};  // cell 64c-1's dx = 1 value for sample 64c, per row
  for (int s0 = 0; s0 < G.Ws; s0 += kGatherStep) {
    const int cl = s0 + (lane & 15);
    const bool ok = row_ok && cl < G.W;
    const int64_t widx = rw + min(cl, G.W - 1);
    const uint64_t wl = ok ? act[widx] : 0ull;
    const uint32_t pl = wprefix[widx];
    const int nstep = min(kGatherStep, G.Ws - s0);
    const bool carried = carry[0] != 0.f || carry[1] != 0.f || carry[2] != 0.f || carry[3] != 0.f;
    if (__ballot(wl != 0ull) == 0ull && !carried) {  // no active cell touches these samples
      for (int x = 64 * s0 + lane; x < min(G.nx, 64 * (s0 + nstep)); x += 64) out[x] = 0.f;
#pragma unroll
      for (int q = 0; q < 4; ++q) carry[q] = 0.f;
      continue;
    }
    for (int b0 = 0; b0 < nstep; b0 += kGatherChunks) {
      float own[kGatherChunks][4], nxt[kGatherChunks][4];
#pragma unroll
      for (int cc = 0; cc < kGatherChunks; ++cc) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int src = 16 * q + ((b0 + cc) & 15);
          const uint64_t w =
              (b0 + cc < nstep)
                  ? (((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(wl >> 32), src) << 32) |
                     (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)wl, src))
                  : 0ull;
          own[cc][q] = 0.f;
          nxt[cc][q] = 0.f;
          if (w) {  // uniform
            const uint32_t p = (uint32_t)__builtin_amdgcn_readlane((int)pl, src);
            const bool a = (w >> lane) & 1ull;
            const uint64_t vid = a ? (uint64_t)p + (uint64_t)__popcll(w & below) : 0ull;
            const int dy = 1 - (q & 1), dz = 1 - (q >> 1);
            // corners (dx=0, dx=1) on this sample row: dy = 0 -> (0, 1); dy = 1 -> (3, 2)
            const float2 v =
                *reinterpret_cast<const float2*>(contrib + 8 * vid + 4 * dz + 2 * dy);
            own[cc][q] = a ? (dy ? v.y : v.x) : 0.f;
            nxt[cc][q] = a ? (dy ? v.x : v.y) : 0.f;
          }
        }
      }
#pragma unroll
      for (int cc = 0; cc < kGatherChunks; ++cc) {
        if (b0 + cc >= nstep) break;  // uniform
        float acc = 0.f;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          float prev = __shfl_up(nxt[cc][q], 1);
          if (lane == 0) prev = carry[q];
          carry[q] = __shfl(nxt[cc][q], 63);
          acc += prev;          // cell x-1
          acc += own[cc][q];    // cell x
        }
        const int x = 64 * (s0 + b0 + cc) + lane;
        if (x < G.nx) out[x] = acc;
      }
    }
  }
}

}  // namespace

