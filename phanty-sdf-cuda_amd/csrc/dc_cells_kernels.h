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
#include "dc_decode.h"
#include "dc_internal.h"
#include "dc_kernels.h"

namespace {

using namespace dccell;
using dccommon::find_bucket;
using dccommon::wave_scan_incl;
using dcdecode::decode_cell;
using dcdecode::decode_face;
using dcdecode::kErrCellRange;
using dcdecode::kErrFaceRange;

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

// fn(G, a, i, j, k, face id, list entry) for every face, one thread per face-list entry; each block takes a
// contiguous XCD-aware share of the list (at 512^3 this halves the kernels' HBM fetch: list
// neighbours share stencil rows, and adjacent shares run under one L2). An entry that names no quad edge of its grid is skipped and flagged in *err.
template <typename Fn>
__device__ __forceinline__ void each_face(const DcGridDev* __restrict__ grids, int n_grids,
                                          const int64_t* __restrict__ ft_bounds,
                                          const uint64_t* __restrict__ flist,
                                          const uint64_t* __restrict__ ft,
                                          const uint32_t* __restrict__ pw, uint64_t n,
                                          uint32_t* __restrict__ err, Fn&& fn) {
  uint64_t lo, hi;
  dccommon::block_share(n, lo, hi);
  for (uint64_t f = lo + threadIdx.x; f < hi; f += blockDim.x) {
    const uint64_t entry = flist[f];
    const int64_t word = (int64_t)(entry >> 6);
    const uint64_t below = (1ull << (entry & 63)) - 1ull;
    const uint64_t face = (uint64_t)pw[word] + (uint64_t)__popcll(ft[word] & below);
    const int g = find_bucket(ft_bounds, n_grids, word);
    const DcGridDev G = grids[g];
    int a, i, j, k;
    if (!decode_face(G, entry, a, i, j, k)) {
      atomicOr(err, kErrFaceRange);
      continue;
    }
    fn(G, a, i, j, k, face, entry);
  }
}

// Averaged hermite normal of every face edge, once per face (plain: FP64 in the global frame;
// FP32 mode: FP32 in each cell's frame): normalized sum and the sum's length.
template <typename T>
__global__ void __launch_bounds__(256)
    k_face_normals(const DcGridDev* __restrict__ grids, int n_grids,
                   const int64_t* __restrict__ ft_bounds, const uint64_t* __restrict__ flist,
                   const uint64_t* __restrict__ ft, const uint32_t* __restrict__ pw,
                   DcCounters* __restrict__ ctr, uint64_t capF, T* __restrict__ nrm,
                   float* __restrict__ nrm32) {
  each_face(grids, n_grids, ft_bounds, flist, ft, pw, min(ctr->n_faces, capF), &ctr->error,
            [&](const DcGridDev& G, int a, int i, int j, int k, uint64_t face, uint64_t) {
              if (face >= capF) return;
              T n[3];
              // the four members from their shared corner block
              if (a == 0)
                group_sum_block<0>(G, i, j, k, n);
              else if (a == 1)
                group_sum_block<1>(G, i, j, k, n);
              else
                group_sum_block<2>(G, i, j, k, n);
              const T len = tsqrt(sq3(n[0], n[1], n[2]));
              normalize3(n);
              nrm[4 * face + 0] = n[0];
              nrm[4 * face + 1] = n[1];
              nrm[4 * face + 2] = n[2];
              nrm[4 * face + 3] = len;
              if (nrm32)  // FP32 mode: the same normals rounded, for the FP32 solve and adjoint
                reinterpret_cast<float4*>(nrm32)[face] =
                    make_float4((float)n[0], (float)n[1], (float)n[2], (float)len);
            });
}

// ---------------------------------------------------------------------------------------------
// Vertices
// ---------------------------------------------------------------------------------------------
// One thread per active cell (persistent grid-stride over the device-side count). In FP32 mode a
// cell that needs FP64 is appended to fblist (order irrelevant: results are per vertex), and the
// adjoint's per-cell row kinds are recorded in rows (12 per vertex) when it is not null. The FP64
// instance is held to 3 waves per SIMD (166 VGPRs, no spills): 110 us at 512^3 against 127 at the
// 2 waves its unbounded register use gives; the FP32 one fits 4.
template <int PREC>
__global__ void __launch_bounds__(256, PREC == DC_PREC_FP32 ? 4 : 3)
    k_qef(const DcGridDev* __restrict__ grids, int n_grids, const int64_t* __restrict__ act_bounds,
          const uint64_t* __restrict__ act, const uint64_t* __restrict__ alist,
          DcCounters* __restrict__ ctr, uint64_t cap, double* __restrict__ v64,
          float* __restrict__ v32, uint32_t* __restrict__ flags, uint32_t* __restrict__ fblist,
          const uint64_t* __restrict__ ft, const uint32_t* __restrict__ pw,
          const void* __restrict__ nrm, uint64_t capF, float* __restrict__ qstate,
          uint32_t* __restrict__ rows) {
  const uint64_t n = min(ctr->n_active, cap);
  const uint64_t nf = min(ctr->n_faces, capF);
  // nrm: the FP64 table (FP64 instance) or its FP32 copy (FP32 instance)
  const FaceNormals<double> fn64{ft, pw, reinterpret_cast<const double*>(nrm), nullptr, nf};
  const FaceNormals<float> fn32{ft, pw, nullptr, reinterpret_cast<const float*>(nrm), nf};
  // FP32 mode: each thread's row kinds gathered in its LDS column (e * 256 + thread), then written
  // as three 16-byte stores, so a wave writes whole lines
  __shared__ uint32_t s_kinds[PREC == DC_PREC_FP32 ? 12 * kKindStride : 1];
  // FP32 solve: contiguous XCD-aware shares (106 against 116 us at 512^3, half the HBM fetch);
  // the FP64 solve is faster on the plain grid stride (110 against 120 us)
  uint64_t lo, hi, step;
  if (PREC == DC_PREC_FP32) {
    dccommon::block_share(n, lo, hi);
    lo += threadIdx.x;
    step = blockDim.x;
  } else {
    lo = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    hi = n;
    step = (uint64_t)gridDim.x * blockDim.x;
  }
  for (uint64_t v = lo; v < hi; v += step) {
    const uint64_t entry = alist[v];
    const int g = find_bucket(act_bounds, n_grids, (int64_t)(entry >> 6));
    const DcGridDev G = grids[g];
    int i, j, k;
    if (!decode_cell(G, entry, i, j, k)) {
      atomicOr(&ctr->error, kErrCellRange);
      continue;
    }
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
      uint32_t* kinds = rows ? s_kinds + threadIdx.x : nullptr;
      const bool ok = solve_cell_fp32(G, act, fn32, i, j, k, xl,
                                      qstate ? qstate + kQefState * v : nullptr, kinds);
      if (kinds) {  // a non-crossing edge's entry is stale LDS content, never read (the readers
                    // mask with the cell's crossing edges)
        uint4* o = reinterpret_cast<uint4*>(rows + 12 * v);
#pragma unroll
        for (int q = 0; q < 3; ++q)
          o[q] = make_uint4(kinds[(4 * q) * kKindStride], kinds[(4 * q + 1) * kKindStride],
                            kinds[(4 * q + 2) * kKindStride], kinds[(4 * q + 3) * kKindStride]);
      }
      if (ok) {
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

// FP32 mode: the listed cells solved in FP64 (the plain arithmetic, with the FP64 face table). The
// solve's state of list entry t < cap_state goes to state + kQefState * t for the adjoint.
// 16 lanes per cell: lane e computes the row of edge e (an averaged normal takes up to 4 raw
// normals), then every lane of the group sums the rows in ascending edge order through shuffles
// (the serial solve's operation order) and solves; the group's first lane writes the vertex.
// The lanes of a group share their cell and so their loop trip counts: the shuffles only read
// lanes of their own, active, group.
constexpr int kFbLanes = 16;

__global__ void __launch_bounds__(256)
    k_qef_fallback(const DcGridDev* __restrict__ grids, int n_grids,
                   const int64_t* __restrict__ act_bounds, const uint64_t* __restrict__ act,
                   const uint64_t* __restrict__ alist, const DcCounters* __restrict__ ctr,
                   uint64_t cap, const uint32_t* __restrict__ fblist, float* __restrict__ v32,
                   double* __restrict__ state, uint64_t cap_state, const uint64_t* __restrict__ ft,
                   const uint32_t* __restrict__ pw, const double* __restrict__ nrm64,
                   uint64_t capF) {
  const uint64_t n = min(ctr->fallback, cap);
  const FaceNormals<double> fn64{ft, pw, nrm64, nullptr, min(ctr->n_faces, capF)};
  const int lane = threadIdx.x & 63;
  const int sub = lane & (kFbLanes - 1);
  const int base = lane - sub;
  const uint64_t groups = (uint64_t)gridDim.x * blockDim.x / kFbLanes;
  for (uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kFbLanes; t < n;
       t += groups) {
    const uint32_t v = fblist[t];
    const uint64_t entry = alist[v];
    const int g = find_bucket(act_bounds, n_grids, (int64_t)(entry >> 6));
    const DcGridDev G = grids[g];
    int i, j, k;
    if (!decode_cell(G, entry, i, j, k)) continue;  // k_qef flagged it (the group skips together)
    Cell<double> own;
    load_cell(G, i, j, k, own);
    const uint32_t cross = crossing_mask(own);
    double p[3] = {0.0, 0.0, 0.0}, nf[3] = {0.0, 0.0, 0.0};
    if (sub < 12 && ((cross >> sub) & 1u)) fp64_row(G, act, fn64, i, j, k, own, sub, p, nf);
    double x[3];
    solve_rows_fp64(
        cross,
        [&](int e, double pe[3], double ne[3]) {
          const int src = base + e;
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            pe[d] = __shfl(p[d], src);
            ne[d] = __shfl(nf[d], src);
          }
        },
        x, (sub == 0 && state && t < cap_state) ? state + kQefState * t : nullptr);
    if (sub == 0) {
      v32[3 * (uint64_t)v + 0] = (float)x[0];
      v32[3 * (uint64_t)v + 1] = (float)x[1];
      v32[3 * (uint64_t)v + 2] = (float)x[2];
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Quads (contouring.cpp:607-729)
// ---------------------------------------------------------------------------------------------
// contouring.cpp:141-207 gradientAt: central differences with clamped neighbours, normalized.
// The six central-difference samples of grid point (i,j,k) (clamped at the border): x+, x-, y+,
// y-, z+, z-. Loaded unconditionally, so a caller can have them in flight with other loads.
__device__ __forceinline__ void gradient_samples(const DcGridDev& G, int i, int j, int k,
                                                 float f[6]) {
  const int im = max(i - 1, 0), ip = min(i + 1, G.nx - 1);
  const int jm = max(j - 1, 0), jp = min(j + 1, G.ny - 1);
  const int km = max(k - 1, 0), kp = min(k + 1, G.nz - 1);
  f[0] = *sample_ptr(G, ip, j, k);
  f[1] = *sample_ptr(G, im, j, k);
  f[2] = *sample_ptr(G, i, jp, k);
  f[3] = *sample_ptr(G, i, jm, k);
  f[4] = *sample_ptr(G, i, j, kp);
  f[5] = *sample_ptr(G, i, j, km);
}

// contouring.cpp:141-207 (gradientAt): the normalized central-difference gradient at grid point
// (i,j,k) from its samples f (gradient_samples); a component whose spacing is below 1e-12 is 0.
__device__ __forceinline__ void grid_gradient_from(const DcGridDev& G, int i, int j, int k,
                                                   const float f[6], double out[3]) {
  const int im = max(i - 1, 0), ip = min(i + 1, G.nx - 1);
  const int jm = max(j - 1, 0), jp = min(j + 1, G.ny - 1);
  const int km = max(k - 1, 0), kp = min(k + 1, G.nz - 1);
  const double dx = coordd(G.ox, ip, G.vx) - coordd(G.ox, im, G.vx);
  const double dy = coordd(G.oy, jp, G.vy) - coordd(G.oy, jm, G.vy);
  const double dz = coordd(G.oz, kp, G.vz) - coordd(G.oz, km, G.vz);
  out[0] = fabs(dx) > 1e-12 ? ((double)f[0] - (double)f[1]) / dx : 0.0;
  out[1] = fabs(dy) > 1e-12 ? ((double)f[2] - (double)f[3]) / dy : 0.0;
  out[2] = fabs(dz) > 1e-12 ? ((double)f[4] - (double)f[5]) / dz : 0.0;
  const double nrm = sqrt(sq3(out[0], out[1], out[2]));
  if (nrm > 0.0) {
    out[0] = out[0] / nrm;
    out[1] = out[1] / nrm;
    out[2] = out[2] / nrm;
  }
}

// Loads of a few consecutive values at the alignment of their element (global memory takes
// dword-aligned multi-dword loads): one instruction where element loads would take several.
typedef float f3a4 __attribute__((ext_vector_type(3), aligned(4)));
typedef float f4a4 __attribute__((ext_vector_type(4), aligned(4)));
typedef double d2a8 __attribute__((ext_vector_type(2), aligned(8)));
typedef unsigned long long u64x2a8 __attribute__((ext_vector_type(2), aligned(8)));
typedef unsigned int u32x2a4 __attribute__((ext_vector_type(2), aligned(4)));
template <typename V, typename E>
__device__ __forceinline__ V load_vec(const E* p) {
  return *(const __attribute__((address_space(1))) V*)(p);  // a global-memory load
}

// gradient_samples at both endpoints of the axis-A quad edge at (i,j,k) (f0 at (i,j,k), f1 one step
// along A), by rows: a quad edge's four cells lie inside the grid, so only the along-axis
// neighbours are clamped, and the samples of both stencils come from 5 (x edge) or 8 (y, z edge)
// loads instead of 12. An x edge on a grid under 4 samples wide takes the element loads.
template <int A>
__device__ __forceinline__ void edge_gradient_samples(const DcGridDev& G, int i, int j, int k,
                                                      float f0[6], float f1[6]) {
  if constexpr (A == 0) {
    // the window starts at i-1 inside, at 0 for i = 0 and at i-2 for i = nx-2 (an x quad edge
    // has i <= nx-2); with nx >= 4 the two border cases exclude each other. Narrower grids take
    // the element loads.
    const bool wide = G.nx >= 4;
    const int xs = wide ? min(max(i - 1, 0), G.nx - 4) : 0;
    const bool left = i == 0, right = i + 2 > G.nx - 1;
    const gfloat yp = sample_ptr(G, i, j + 1, k), ym = sample_ptr(G, i, j - 1, k);
    const gfloat zp = sample_ptr(G, i, j, k + 1), zm = sample_ptr(G, i, j, k - 1);
    f0[2] = yp[0];
    f1[2] = yp[1];
    f0[3] = ym[0];
    f1[3] = ym[1];
    f0[4] = zp[0];
    f1[4] = zp[1];
    f0[5] = zm[0];
    f1[5] = zm[1];
    if (wide) {
      const f4a4 w = load_vec<f4a4>(&sample_ptr(G, xs, j, k)[0]);
      f0[0] = left ? w.y : (right ? w.w : w.z);  // i+1
      f0[1] = right ? w.y : w.x;                 // max(i-1, 0)
      f1[0] = left ? w.z : w.w;                  // min(i+2, nx-1)
      f1[1] = left ? w.x : (right ? w.z : w.y);  // i
    } else {
      const gfloat r = sample_ptr(G, 0, j, k);
      f0[0] = r[i + 1];
      f0[1] = r[max(i - 1, 0)];
      f1[0] = r[min(i + 2, G.nx - 1)];
      f1[1] = r[i];
    }
  } else {
    // rows through both endpoints give the x neighbours and the endpoints themselves
    const int j1 = j + (A == 1), k1 = k + (A == 2);
    const f3a4 r0 = load_vec<f3a4>(&sample_ptr(G, i - 1, j, k)[0]);
    const f3a4 r1 = load_vec<f3a4>(&sample_ptr(G, i - 1, j1, k1)[0]);
    f0[0] = r0.z;
    f0[1] = r0.x;
    f1[0] = r1.z;
    f1[1] = r1.x;
    if constexpr (A == 1) {
      const int jm0 = max(j - 1, 0), jp1 = min(j + 2, G.ny - 1);
      f0[2] = r1.y;                         // (i, j+1, k)
      f0[3] = *sample_ptr(G, i, jm0, k);
      f1[2] = *sample_ptr(G, i, jp1, k);
      f1[3] = r0.y;                         // (i, j, k)
      f0[4] = *sample_ptr(G, i, j, k + 1);
      f0[5] = *sample_ptr(G, i, j, k - 1);
      f1[4] = *sample_ptr(G, i, j1, k + 1);
      f1[5] = *sample_ptr(G, i, j1, k - 1);
    } else {
      const int km0 = max(k - 1, 0), kp1 = min(k + 2, G.nz - 1);
      f0[2] = *sample_ptr(G, i, j + 1, k);
      f0[3] = *sample_ptr(G, i, j - 1, k);
      f1[2] = *sample_ptr(G, i, j + 1, k1);
      f1[3] = *sample_ptr(G, i, j - 1, k1);
      f0[4] = r1.y;                         // (i, j, k+1)
      f0[5] = *sample_ptr(G, i, j, km0);
      f1[4] = *sample_ptr(G, i, j, kp1);
      f1[5] = r0.y;                         // (i, j, k)
    }
  }
}

// Vertex ids (grid-local) of the active cells (i-1, j, k) and (i, j, k), i >= 1, from one load of
// the two act words at and after cell i-1's and one of their prefixes. The act and prefix buffers
// hold one word past their end, so the second word can always be loaded; it is used only when
// cell i starts a new word (then it is in the same row).
__device__ __forceinline__ void vertex_id_pair(const DcGridDev& G,
                                               const uint64_t* __restrict__ act,
                                               const uint32_t* __restrict__ wprefix,
                                               uint32_t vbase, int i, int j, int k, int32_t& lo,
                                               int32_t& hi) {
  const int64_t w = G.act_off + ((int64_t)k * G.cy + j) * G.W + ((i - 1) >> 6);
  const u64x2a8 a = load_vec<u64x2a8>(act + w);
  const u32x2a4 p = load_vec<u32x2a4>(wprefix + w);
  lo = (int32_t)(p.x - vbase + (uint32_t)__popcll(a.x & ((1ull << ((i - 1) & 63)) - 1ull)));
  hi = (i & 63) ? (int32_t)(p.x - vbase + (uint32_t)__popcll(a.x & ((1ull << (i & 63)) - 1ull)))
                : (int32_t)(p.y - vbase);
}

// Vertex id (grid-local) of active cell (i,j,k): rank in (k,j,i) order.
__device__ __forceinline__ int32_t vertex_id(const DcGridDev& G, const uint64_t* __restrict__ act,
                                             const uint32_t* __restrict__ wprefix, uint32_t vbase,
                                             int i, int j, int k) {
  const int64_t w = G.act_off + ((int64_t)k * G.cy + j) * G.W + (i >> 6);
  const uint64_t below = (1ull << (i & 63)) - 1ull;
  return (int32_t)(wprefix[w] - vbase + (uint32_t)__popcll(act[w] & below));
}

// contouring.cpp:607-729: the quad (c0, c1, c2, c3) turned to (c0, c3, c2, c1) when its normal
// (pb - pa) x (pd - pa), or (pc - pa) x (pd - pa) when that has squared length below 1e-24, points
// against gn, the sum of the edge endpoints' normalized gradients. Returns q . gn (0 when either
// vector is zero and the quad is left as it is); *q1sq receives the squared length of the first
// normal (the one the 1e-24 test is made on).
__device__ __forceinline__ double orient_quad(const double pa[3], const double pb[3],
                                              const double pc[3], const double pd[3], double gn0,
                                              double gn1, double gn2, int& c1, int& c3,
                                              double* q1sq = nullptr) {
  const double ba0 = pb[0] - pa[0], ba1 = pb[1] - pa[1], ba2 = pb[2] - pa[2];
  const double da0 = pd[0] - pa[0], da1 = pd[1] - pa[1], da2 = pd[2] - pa[2];
  double q0 = ba1 * da2 - ba2 * da1;
  double q1 = ba2 * da0 - ba0 * da2;
  double q2 = ba0 * da1 - ba1 * da0;
  const double first = sq3(q0, q1, q2);
  if (q1sq) *q1sq = first;
  if (first < 1e-24) {
    const double ca0 = pc[0] - pa[0], ca1 = pc[1] - pa[1], ca2 = pc[2] - pa[2];
    q0 = ca1 * da2 - ca2 * da1;
    q1 = ca2 * da0 - ca0 * da2;
    q2 = ca0 * da1 - ca1 * da0;
  }
  double dot = 0.0;
  if (sq3(gn0, gn1, gn2) > 0 && sq3(q0, q1, q2) > 0) {
    dot = (q0 * gn0 + q1 * gn1) + q2 * gn2;
    if (dot < 0.0) {
      const int t = c1;
      c1 = c3;
      c3 = t;
    }
  }
  return dot;
}

// A quad's four cells (c0..c3 in orient_quad's order) around the axis-a edge at (i, j, k):
// c0 = (u-1, v-1), c1 = (u, v-1), c2 = (u, v), c3 = (u-1, v) over the two axes (u, v) across the
// edge (x: (y, z); y: (x, z); z: (x, y)).
__device__ __forceinline__ void quad_cell(int a, int i, int j, int k, int q, int& x, int& y,
                                          int& z) {
  const int ux = (a != 0), uy = (a == 0);  // unit step of u
  const int vy = (a == 2), vz = (a != 2);  // unit step of v
  const int du = (q == 0 || q == 3), dv = (q == 0 || q == 1);
  x = i - du * ux;
  y = j - du * uy - dv * vy;
  z = k - dv * vz;
}

// FP32 mode's near-tie bound: an FP32 vertex is within kVertexTol * h of the FP64 one (the mode's
// stated tolerance), so a quad normal from FP32 vertices is within about
// 4 kVertexTol h (|b - a| + |c - a| + |d - a|) of the FP64 one. A decision is made again from FP64
// vertices when, with kTieMargin times that error, its dot product could change sign, or the first
// normal's length could fall on the other side of the degenerate-quad test (1e-12).
constexpr double kVertexTol = 1e-4;
constexpr double kTieMargin = 8.0;

// One quad: every load issued before its arithmetic (the stencil rows of both endpoints, the four
// cells' act words and prefixes, their vertices), then the orientation test of contouring.cpp:
// 607-729; FP32 mode lists the quad for k_reorient when the test is within its error bound.
template <int A>
__device__ __forceinline__ void emit_face(const DcGridDev& G, int i, int j, int k, uint64_t pos,
                                          uint64_t entry, const uint64_t* __restrict__ act,
                                          const uint32_t* __restrict__ wprefix,
                                          const double* __restrict__ v64,
                                          const float* __restrict__ v32, int precision,
                                          DcCounters* __restrict__ ctr, uint64_t capF,
                                          int32_t* __restrict__ faces,
                                          uint64_t* __restrict__ reorient) {
  const int e1x = i + (A == 0), e1y = j + (A == 1), e1z = k + (A == 2);  // far endpoint
  float f0[6], f1[6];
  edge_gradient_samples<A>(G, i, j, k, f0, f1);
  const uint32_t vbase = wprefix[G.act_off];
  const int64_t vglob = (int64_t)vbase;  // grid's first vertex in the concatenated arrays
  int c[4];
  if constexpr (A == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int x, y, z;
      quad_cell(A, i, j, k, q, x, y, z);
      c[q] = vertex_id(G, act, wprefix, vbase, x, y, z);
    }
  } else {
    // cells (i-1, i) in two rows: c0 = (i-1, r0), c1 = (i, r0), c2 = (i, r1), c3 = (i-1, r1)
    const int jr0 = A == 2 ? j - 1 : j, kr0 = A == 1 ? k - 1 : k;
    vertex_id_pair(G, act, wprefix, vbase, i, jr0, kr0, c[0], c[1]);
    vertex_id_pair(G, act, wprefix, vbase, i, j, k, c[3], c[2]);
  }
  double pa[3], pb[3], pc[3], pd[3];
  double* dst[4] = {pa, pb, pc, pd};
  if (precision == DC_PREC_FP64) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double* v = v64 + 3 * (vglob + c[q]);
      const d2a8 xy = load_vec<d2a8>(v);
      dst[q][0] = xy.x;
      dst[q][1] = xy.y;
      dst[q][2] = v[2];
    }
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const f3a4 xyz = load_vec<f3a4>(v32 + 3 * (vglob + c[q]));
      dst[q][0] = (double)xyz.x;
      dst[q][1] = (double)xyz.y;
      dst[q][2] = (double)xyz.z;
    }
  }
  double g0[3], g1[3];
  grid_gradient_from(G, i, j, k, f0, g0);
  grid_gradient_from(G, e1x, e1y, e1z, f1, g1);
  const double gn0 = g0[0] + g1[0], gn1 = g0[1] + g1[1], gn2 = g0[2] + g1[2];
  int c1 = c[1], c3 = c[3];
  double q1sq = 0.0;
  const double dot = orient_quad(pa, pb, pc, pd, gn0, gn1, gn2, c1, c3, &q1sq);
  if (precision != DC_PREC_FP64 && reorient) {
    const double ba = tsqrt(sq3(pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]));
    const double ca = tsqrt(sq3(pc[0] - pa[0], pc[1] - pa[1], pc[2] - pa[2]));
    const double da = tsqrt(sq3(pd[0] - pa[0], pd[1] - pa[1], pd[2] - pa[2]));
    const double gl = tsqrt(sq3(gn0, gn1, gn2));
    const double err =
        kTieMargin * 4.0 * kVertexTol * fmin(fmin(G.vx, G.vy), G.vz) * (ba + ca + da) + 2e-12;
    if (gl > 0.0 && (tsqrt(q1sq) <= 1e-12 + err || fabs(dot) <= err * gl)) {
      const uint64_t slot = atomicAdd((unsigned long long*)&ctr->n_reorient, 1ull);
      if (slot < capF) reorient[slot] = entry;
    }
  }
  reinterpret_cast<int4*>(faces)[pos] = make_int4(c[0], c1, c[2], c3);
}

__global__ void __launch_bounds__(256)
    k_emit(const DcGridDev* __restrict__ grids, int n_grids, const int64_t* __restrict__ ft_bounds,
           const uint64_t* __restrict__ flist, const uint64_t* __restrict__ ft,
           const uint32_t* __restrict__ pw, const uint64_t* __restrict__ act,
           const uint32_t* __restrict__ wprefix, const double* __restrict__ v64,
           const float* __restrict__ v32, int precision, DcCounters* __restrict__ ctr,
           uint64_t capA, uint64_t capF, int32_t* __restrict__ faces,
           uint64_t* __restrict__ reorient) {
  // an overflowed run is repeated by the host; never index past the buffers
  if (ctr->n_active > capA || ctr->n_faces > capF) return;
  each_face(grids, n_grids, ft_bounds, flist, ft, pw, ctr->n_faces, &ctr->error,
            [&](const DcGridDev& G, int a, int i, int j, int k, uint64_t pos, uint64_t entry) {
    if (a == 0)
      emit_face<0>(G, i, j, k, pos, entry, act, wprefix, v64, v32, precision, ctr, capF, faces,
                   reorient);
    else if (a == 1)
      emit_face<1>(G, i, j, k, pos, entry, act, wprefix, v64, v32, precision, ctr, capF, faces,
                   reorient);
    else
      emit_face<2>(G, i, j, k, pos, entry, act, wprefix, v64, v32, precision, ctr, capF, faces,
                   reorient);
  });
}

// FP32 mode: the quads k_emit listed as near ties, oriented again from FP64 vertices of their four
// cells, solved here with the plain mode's arithmetic (this unit has no FMA contraction; averaged
// normals from the FP64 face table, as in the plain solve). One wave per quad:
// lanes 16q..16q+15 solve cell c_q, one crossing row per lane summed in ascending edge order (the
// serial solve's operation order); lane 0 then orients the quad like the plain mode's k_emit. The
// result equals the plain mode's quad.
__global__ void __launch_bounds__(256)
    k_reorient(const DcGridDev* __restrict__ grids, int n_grids,
               const int64_t* __restrict__ ft_bounds, const uint64_t* __restrict__ ft,
               const uint32_t* __restrict__ pw, const uint64_t* __restrict__ act,
               const uint32_t* __restrict__ wprefix, const DcCounters* __restrict__ ctr,
               uint64_t capF, const uint64_t* __restrict__ reorient,
               const double* __restrict__ nrm64, int32_t* __restrict__ faces) {
  const uint64_t n = min(ctr->n_reorient, capF);
  const FaceNormals<double> fn64{ft, pw, nrm64, nullptr, min(ctr->n_faces, capF)};
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4, sub = lane & 15;
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x >> 6);
  for (uint64_t t = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); t < n;
       t += waves) {
    const uint64_t entry = reorient[t];
    const int64_t word = (int64_t)(entry >> 6);
    const uint64_t pos = (uint64_t)pw[word] + (uint64_t)__popcll(ft[word] & ((1ull << (entry & 63)) - 1ull));
    const DcGridDev G = grids[find_bucket(ft_bounds, n_grids, word)];
    int a, i, j, k;
    if (!decode_face(G, entry, a, i, j, k) || pos >= capF) continue;  // uniform per wave
    int x, y, z;
    quad_cell(a, i, j, k, q, x, y, z);
    Cell<double> own;
    load_cell(G, x, y, z, own);
    const uint32_t cross = crossing_mask(own);
    double p[3] = {0.0, 0.0, 0.0}, nf[3] = {0.0, 0.0, 0.0};
    if (sub < 12 && ((cross >> sub) & 1u)) fp64_row(G, act, fn64, x, y, z, own, sub, p, nf);
    double v[3];
    solve_rows_fp64(
        cross,
        [&](int e, double pe[3], double ne[3]) {
          const int src = 16 * q + e;
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            pe[d] = __shfl(p[d], src);
            ne[d] = __shfl(nf[d], src);
          }
        },
        v);
    double P[4][3];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int d = 0; d < 3; ++d) P[m][d] = __shfl(v[d], 16 * m);
    if (lane == 0) {
      const int e1x = i + (a == 0), e1y = j + (a == 1), e1z = k + (a == 2);
      float f0[6], f1[6];
      gradient_samples(G, i, j, k, f0);
      gradient_samples(G, e1x, e1y, e1z, f1);
      double g0[3], g1[3];
      grid_gradient_from(G, i, j, k, f0, g0);
      grid_gradient_from(G, e1x, e1y, e1z, f1, g1);
      const uint32_t vbase = wprefix[G.act_off];
      int c[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        int cx, cy, cz;
        quad_cell(a, i, j, k, m, cx, cy, cz);
        c[m] = vertex_id(G, act, wprefix, vbase, cx, cy, cz);
      }
      int c1 = c[1], c3 = c[3];
      orient_quad(P[0], P[1], P[2], P[3], g0[0] + g1[0], g0[1] + g1[1], g0[2] + g1[2], c1, c3);
      reinterpret_cast<int4*>(faces)[pos] = make_int4(c[0], c1, c[2], c3);
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Triangles
// ---------------------------------------------------------------------------------------------
// triangulate_v1 (contouring.cpp:88-113) with orient_triangles_for_quad (contouring.cpp:40-86),
// one thread per quad: (a, b, c, d) -> (a, b, c), (a, c, d); the quad normal is
// (pb - pa) x (pd - pa), or (pc - pa) x (pd - pa) when that has squared length < 1e-24; a quad
// whose normal is still that short keeps both triangles; otherwise a triangle whose normal points
// against the quad normal has its last two vertices swapped. Positions are the FP64 vertices
// (plain mode) or the FP32 ones widened. In a batch, quad f's vertex ids are local to its grid:
// goff holds each grid's (first vertex, first face).
__device__ __forceinline__ void tri_sub(const double u[3], const double w[3], double o[3]) {
  o[0] = u[0] - w[0];
  o[1] = u[1] - w[1];
  o[2] = u[2] - w[2];
}
__device__ __forceinline__ void tri_cross(const double u[3], const double w[3], double o[3]) {
  o[0] = u[1] * w[2] - u[2] * w[1];
  o[1] = u[2] * w[0] - u[0] * w[2];
  o[2] = u[0] * w[1] - u[1] * w[0];
}
__device__ __forceinline__ double tri_dot(const double u[3], const double w[3]) {
  return (u[0] * w[0] + u[1] * w[1]) + u[2] * w[2];
}

__global__ void __launch_bounds__(256)
    k_triangulate(const int4* __restrict__ faces, uint64_t nf, int n_grids,
                  const uint32_t* __restrict__ goff, const double* __restrict__ v64,
                  const float* __restrict__ v32, int use64, int32_t* __restrict__ tris) {
  for (uint64_t f = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; f < nf;
       f += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t vb = 0;
    if (n_grids > 1) {  // last grid whose first face is <= f
      int lo = 0, hi = n_grids - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((uint64_t)goff[2 * mid + 1] <= f)
          lo = mid;
        else
          hi = mid - 1;
      }
      vb = goff[2 * lo];
    }
    const int4 q = faces[f];
    auto P = [&](int32_t id, double o[3]) {
      const uint64_t b = 3 * (vb + (uint64_t)id);
      if (use64) {
        o[0] = v64[b];
        o[1] = v64[b + 1];
        o[2] = v64[b + 2];
      } else {
        o[0] = (double)v32[b];
        o[1] = (double)v32[b + 1];
        o[2] = (double)v32[b + 2];
      }
    };
    int32_t t[2][3] = {{q.x, q.y, q.z}, {q.x, q.z, q.w}};
    double pa[3], pb[3], pc[3], pd[3], u[3], w[3], qn[3];
    P(q.x, pa);
    P(q.y, pb);
    P(q.z, pc);
    P(q.w, pd);
    tri_sub(pb, pa, u);
    tri_sub(pd, pa, w);
    tri_cross(u, w, qn);
    if (tri_dot(qn, qn) < 1e-24) {
      tri_sub(pc, pa, u);
      tri_cross(u, w, qn);
    }
    if (tri_dot(qn, qn) >= 1e-24) {
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        double p1[3], p2[3], p3[3], tn[3];
        P(t[r][0], p1);
        P(t[r][1], p2);
        P(t[r][2], p3);
        tri_sub(p2, p1, u);
        tri_sub(p3, p1, w);
        tri_cross(u, w, tn);
        if (tri_dot(tn, qn) < 0.0) {
          const int32_t x = t[r][1];
          t[r][1] = t[r][2];
          t[r][2] = x;
        }
      }
    }
    int32_t* o = tris + 6 * f;
    o[0] = t[0][0];
    o[1] = t[0][1];
    o[2] = t[0][2];
    o[3] = t[1][0];
    o[4] = t[1][1];
    o[5] = t[1][2];
  }
}

// ---------------------------------------------------------------------------------------------
// Adjoint
// ---------------------------------------------------------------------------------------------
// Cells processed: every active cell (list == nullptr; with skip_flagged the FP32 kernel leaves
// the FP64 cells to the list launch) or the listed ones (the FP64 cells of FP32 mode). The cell's 8 corner adjoints go to contrib[8v..8v+7] (phase 1 writes,
// phase 2 adds). The kinds of the cell's rows, rows[12v..12v+11], come from the FP32 forward in
// FP32 mode; in plain mode phase 1 records them (gathered in LDS, then written as three 16-byte
// stores).
// kGiven (phase 1, FP32 mode): the forward's solve state is used, at state + kQefState * v (FP32
// cells) or state + kQefState * t (FP64 cells: the fallback solve's, by list index).
// FP32 instances fit 4 waves per SIMD without spilling; FP64 ones keep the default bound.
// S: the type face_adjoint stored the faces' adjoints in (phase 2).
template <int PHASE, typename T, typename S, bool kGiven>
__global__ void __launch_bounds__(256, sizeof(T) == 4 ? 4 : 1)
    k_bwd_cell(const DcGridDev* __restrict__ grids, int n_grids,
               const int64_t* __restrict__ act_bounds, const uint64_t* __restrict__ act,
               const uint32_t* __restrict__ wprefix, const uint64_t* __restrict__ alist,
               const uint32_t* __restrict__ flags, int skip_flagged,
               const uint32_t* __restrict__ list, const DcCounters* __restrict__ ctr, uint64_t cap, const float* __restrict__ gv,
               const uint64_t* __restrict__ ft, const uint32_t* __restrict__ pw,
               const double* __restrict__ nrm64, const float* __restrict__ nrm32, uint64_t capF,
               const T* __restrict__ state, float* __restrict__ ubar,
               float* __restrict__ ubar_e, uint32_t* __restrict__ rows,
               float* __restrict__ contrib) {
  const uint64_t n = list ? min(ctr->fallback, cap) : min(ctr->n_active, cap);
  const uint64_t nf = min(ctr->n_faces, capF);
  const FaceNormals<T> fn{ft, pw, nrm64, nrm32, nf};
  constexpr bool kRecord = PHASE == 1 && !kGiven;
  __shared__ uint32_t s_kinds[kRecord ? 12 * kKindStride : 1];
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t v = list ? (uint64_t)list[t] : t;
    if (!list && skip_flagged && flags[v]) continue;
    const uint64_t entry = alist[v];
    const int g = find_bucket(act_bounds, n_grids, (int64_t)(entry >> 6));
    const DcGridDev G = grids[g];
    int i, j, k;
    if (!decode_cell(G, entry, i, j, k)) continue;  // the forward flagged it and failed
    T Sb[8];
    float4* out = reinterpret_cast<float4*>(contrib + 8 * v);
    if (PHASE == 1) {
      const float xb[3] = {gv[3 * v], gv[3 * v + 1], gv[3 * v + 2]};
      // plain mode records the kinds (FP32 mode's forward did, for every cell)
      uint32_t* kinds = kRecord && !list ? s_kinds + threadIdx.x : nullptr;
      const T* st = kGiven ? state + kQefState * (sizeof(T) == 4 ? v : t) : nullptr;
      backward_rows<T, kGiven>(G, act, fn, i, j, k, v, xb, st, ubar, ubar_e, rows, kinds, Sb);
      if (kinds) {  // a non-crossing edge's entry is stale LDS content, never read
        uint4* o = reinterpret_cast<uint4*>(rows + 12 * v);
#pragma unroll
        for (int q = 0; q < 3; ++q)
          o[q] = make_uint4(kinds[(4 * q) * kKindStride], kinds[(4 * q + 1) * kKindStride],
                            kinds[(4 * q + 2) * kKindStride], kinds[(4 * q + 3) * kKindStride]);
      }
      out[0] = make_float4((float)Sb[0], (float)Sb[1], (float)Sb[2], (float)Sb[3]);
      out[1] = make_float4((float)Sb[4], (float)Sb[5], (float)Sb[6], (float)Sb[7]);
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) Sb[q] = T(0);
      backward_faces<T, S>(G, act, wprefix, i, j, k, rows + 12 * v, ubar, ubar_e, nf, Sb);
      const float4 o0 = out[0], o1 = out[1];
      out[0] = make_float4(o0.x + (float)Sb[0], o0.y + (float)Sb[1], o0.z + (float)Sb[2],
                           o0.w + (float)Sb[3]);
      out[1] = make_float4(o1.x + (float)Sb[4], o1.y + (float)Sb[5], o1.z + (float)Sb[6],
                           o1.w + (float)Sb[7]);
    }
  }
}

// Edge-parallel form of k_bwd_cell for grids too small to fill the device one cell per thread:
// L lanes per cell (4 or 16, within a row of the wave); lane `sub` takes the cell's crossing rows
// e = sub, sub + L, ... (< 12). The lanes' corner adjoints are summed over the group by an xor
// butterfly (a fixed order: results are reproducible) and the group's first lane writes
// (phase 1) or adds (phase 2) them. Phase 1 uses the forward's solve state and row kinds (FP32
// mode), at state + kQefState * v (FP32 cells) or state + kQefState * t (FP64 cells, by list
// index); every lane forms the cell's RowsAdjoint. The lanes of a group share their cell and so
// their loop trip counts: the shuffles only read lanes of their own, active, group.
template <int PHASE, typename T, typename S, int L>
__global__ void __launch_bounds__(256, sizeof(T) == 4 ? 4 : 1)
    k_bwd_cell_lanes(const DcGridDev* __restrict__ grids, int n_grids,
                     const int64_t* __restrict__ act_bounds, const uint64_t* __restrict__ act,
                     const uint32_t* __restrict__ wprefix, const uint64_t* __restrict__ alist,
                     const uint32_t* __restrict__ flags, int skip_flagged,
                     const uint32_t* __restrict__ list, const DcCounters* __restrict__ ctr,
                     uint64_t cap, const float* __restrict__ gv, const uint64_t* __restrict__ ft,
                     const uint32_t* __restrict__ pw, const double* __restrict__ nrm64,
                     const float* __restrict__ nrm32, uint64_t capF, const T* __restrict__ state,
                     float* __restrict__ ubar, float* __restrict__ ubar_e,
                     uint32_t* __restrict__ rows, float* __restrict__ contrib) {
  static_assert(L == 4 || L == 16, "4 or 16 lanes per cell");
  const uint64_t n = list ? min(ctr->fallback, cap) : min(ctr->n_active, cap);
  const uint64_t nf = min(ctr->n_faces, capF);
  const FaceNormals<T> fn{ft, pw, nrm64, nrm32, nf};
  const int sub = (int)(threadIdx.x & (L - 1));
  const uint64_t groups = (uint64_t)gridDim.x * blockDim.x / L;
  for (uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / L; t < n; t += groups) {
    const uint64_t v = list ? (uint64_t)list[t] : t;
    if (!list && skip_flagged && flags[v]) continue;
    const uint64_t entry = alist[v];
    const int g = find_bucket(act_bounds, n_grids, (int64_t)(entry >> 6));
    const DcGridDev G = grids[g];
    int i, j, k;
    if (!decode_cell(G, entry, i, j, k)) continue;  // the forward flagged it (the group skips)
    // L = 16: the lane's one row, its kind loaded with the cell
    const uint32_t kind16 = L == 16 ? rows[12 * v + (uint64_t)min(sub, 11)] : 0u;
    Cell<T> own;
    load_cell(G, i, j, k, own);
    const uint32_t cross = crossing_mask(own);
    T Sb[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) Sb[q] = T(0);
    if (PHASE == 1) {
      Qef<T> Q;
      T lam[3], V[3][3];
      load_state(state + kQefState * (sizeof(T) == 4 ? v : t), Q, lam, V);
      Q.m = __popc(cross);
      const float xb[3] = {gv[3 * v], gv[3 * v + 1], gv[3 * v + 2]};
      RowsAdjoint<T> R;
      rows_adjoint(Q, lam, V, xb, R);
      if constexpr (L == 16) {
        if (sub < 12 && ((cross >> sub) & 1u))
          backward_row_given(G, act, fn, i, j, k, v, own, R, sub, kind16, ubar, ubar_e, Sb);
      } else {
        for (int e = sub; e < 12; e += L)
          if ((cross >> e) & 1u)
            backward_row_given(G, act, fn, i, j, k, v, own, R, e, rows[12 * v + e], ubar,
                               ubar_e, Sb);
      }
    } else {
      if constexpr (L == 16) {
        if (sub < 12 && ((cross >> sub) & 1u) && kind16 != kRowOwn)
          backward_face_row<T, S>(G, act, wprefix, i, j, k, own, sub, kind16, ubar, ubar_e, nf,
                                  Sb);
      } else {
#pragma unroll 1
        for (int e = sub; e < 12; e += L) {
          if (!((cross >> e) & 1u)) continue;
          const uint32_t kind = rows[12 * v + e];
          if (kind != kRowOwn)
            backward_face_row<T, S>(G, act, wprefix, i, j, k, own, e, kind, ubar, ubar_e, nf, Sb);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q)
#pragma unroll
      for (int off = L / 2; off >= 1; off >>= 1) Sb[q] += __shfl_xor(Sb[q], off, L);
    if (sub == 0) {
      float4* out = reinterpret_cast<float4*>(contrib + 8 * v);
      if (PHASE == 1) {
        out[0] = make_float4((float)Sb[0], (float)Sb[1], (float)Sb[2], (float)Sb[3]);
        out[1] = make_float4((float)Sb[4], (float)Sb[5], (float)Sb[6], (float)Sb[7]);
      } else {
        const float4 o0 = out[0], o1 = out[1];
        out[0] = make_float4(o0.x + (float)Sb[0], o0.y + (float)Sb[1], o0.z + (float)Sb[2],
                             o0.w + (float)Sb[3]);
        out[1] = make_float4(o1.x + (float)Sb[4], o1.y + (float)Sb[5], o1.z + (float)Sb[6],
                             o1.w + (float)Sb[7]);
      }
    }
  }
}

// Between the adjoint phases: each face's averaged-normal adjoint, once per face (face_adjoint),
// stored as double (plain mode) or float (FP32 mode).
template <typename S>
__global__ void __launch_bounds__(256)
    k_bwd_face_adjoint(const DcCounters* __restrict__ ctr, uint64_t capF,
                       const double* __restrict__ nrm64, const float* __restrict__ nrm32,
                       float* __restrict__ ubar) {
  const uint64_t nf = min(ctr->n_faces, capF);
  const FaceNormals<double> fn{nullptr, nullptr, nrm64, nrm32, nf};
  for (uint64_t f = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; f < nf;
       f += (uint64_t)gridDim.x * blockDim.x)
    face_adjoint<S>(fn, f, ubar);
}

// Corner index (contouring.cpp:232-241 numbering) at offset (dx,dy,dz) in {0,1}^3.
__device__ __forceinline__ int corner_at(int dx, int dy, int dz) {
  return 4 * dz + (dy ? (dx ? 2 : 3) : dx);
}

// Act word and vertex prefix for lane 16q + c of the gather: chunk s0 + c of cell row
// (y - 1 + (q & 1), z - 1 + (q >> 1)); the word is 0 outside the grid. The index is clamped so that
// the loads are unconditional.
__device__ __forceinline__ void gather_words(const DcGridDev& G, const uint64_t* __restrict__ act,
                                             const uint32_t* __restrict__ wprefix, int y, int z,
                                             int s0, int lane, uint64_t& wl, uint32_t& pl) {
  const int lq = lane >> 4;
  const int jy = y - 1 + (lq & 1), kz = z - 1 + (lq >> 1);
  const int cl = s0 + (lane & 15);
  const bool ok = jy >= 0 && jy < G.cy && kz >= 0 && kz < G.cz && cl < G.W;
  const int64_t widx = G.act_off +
                       ((int64_t)min(max(kz, 0), G.cz - 1) * G.cy + min(max(jy, 0), G.cy - 1)) *
                           G.W +
                       min(cl, G.W - 1);
  const uint64_t w = act[widx];
  pl = wprefix[widx];
  wl = ok ? w : 0ull;
}

// dL/dphi. One wave per group of R consecutive sample rows (y fastest, then z). Sample x of a row
// is the sum of the corner contributions of the active cells (x-1..x, y-1..y, z-1..z), and every
// sample is written. Per step of up to 16 chunks of 64 samples:
//  - the act words and vertex prefixes of each row's 4 cell rows (q bit 0: j = y, bit 1: k = z)
//    are loaded together for all R rows (lane = 16q + chunk);
//  - the active cells of all R rows are listed in LDS (row, then q, then x), and one float2 load
//    per 64 listed cells fetches each cell's two corner values on its sample row (dx = 0 and 1
//    are adjacent in the corner numbering); all of them are issued before the first sum;
//  - row by row, the values are added into the wave's LDS row buffer at samples x and x+1, one q
//    at a time, so no two lanes of one add touch the same sample; a wave's LDS operations
//    complete in issue order, so every sum is formed in the same order on every run. The row is
//    then stored; a row with no active cell (and nothing carried in) is zero-filled directly. No
//    global load follows a store, so the stores drain only at the end of the wave;
//  - the sample one past the step carries into the next step.
// When a step lists more than kGatherCap cells, the rows are instead summed one at a time, in
// passes of kGatherCap cells.
// LDS per wave: a row buffer of lds_row floats (lds_row >= 64 * chunks per step + 1), then
// kGatherCap cell entries (vertex id, (row << 14) | (q << 12) | sample).
constexpr int kGatherStep = 16;                 // chunks per step
constexpr int kGatherLoads = kGatherCap / 64;   // contribution loads in flight per lane

__device__ __forceinline__ void lds_add(float* p, float v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lists this lane's active cells whose list index k (from `first`) falls in [lo, hi).
__device__ __forceinline__ void list_cells(uint2* __restrict__ cells, uint64_t w, uint32_t p,
                                           uint32_t first, uint32_t lo, uint32_t hi,
                                           uint32_t meta) {
  uint32_t k = first, n = 0;
  while (w && k < hi) {
    const int bit = __builtin_ctzll(w);
    w &= w - 1ull;
    if (k >= lo) cells[k - lo] = make_uint2(p + n, meta + (uint32_t)bit);
    ++k;
    ++n;
  }
}

// One float2 of corner values for listed cell d: (value at sample x, value at sample x+1).
__device__ __forceinline__ float2 cell_pair(const float* __restrict__ contrib, uint2 d) {
  const int q = (int)((d.y >> 12) & 3u);
  const int dy = 1 - (q & 1), dz = 1 - (q >> 1);
  // corners (dx=0, dx=1) on this sample row: dy = 0 -> (0, 1); dy = 1 -> (3, 2)
  const float2 v =
      *reinterpret_cast<const float2*>(contrib + 8 * (uint64_t)d.x + 4 * dz + 2 * dy);
  return dy ? make_float2(v.y, v.x) : v;
}

// Adds the listed cells of row r (lanes with `on`) into the row buffer, one q at a time.
__device__ __forceinline__ void add_cells(float* __restrict__ buf, uint2 d, float2 v, bool on,
                                          int r) {
  const int q = (int)((d.y >> 12) & 3u);
  const bool mine = on && (int)(d.y >> 14) == r;
  float* b = buf + (int)(d.y & 4095u);
#pragma unroll
  for (int qq = 0; qq < 4; ++qq) {
    if (mine && q == qq) {
      lds_add(b, v.x);
      lds_add(b + 1, v.y);
    }
  }
}

// Gradient stores are non-temporal (streamed to HBM): the dense gradient is never read back by
// the library, and cached, its dirty lines were written back during the next extraction's
// sign pass (at 512^3: sign pass 118 -> 98 us after an adjoint, the gather 155 -> 150 us).
__device__ __forceinline__ void gstore(float* p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void gstore(float4* p, float4 v) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  __builtin_nontemporal_store(f4{v.x, v.y, v.z, v.w}, reinterpret_cast<f4*>(p));
}

template <int R>
__global__ void __launch_bounds__(256)
    k_bwd_gather(const DcGridDev* __restrict__ grids, int n_grids,
                 const int64_t* __restrict__ grp_bounds, int64_t n_grp, int lds_row,
                 const uint64_t* __restrict__ act, const uint32_t* __restrict__ wprefix,
                 const float* __restrict__ contrib, float* __restrict__ grad) {
  extern __shared__ float s_gather[];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  float* buf = s_gather + (size_t)wv * (lds_row + 2 * kGatherCap);
  uint2* cells = reinterpret_cast<uint2*>(buf + lds_row);
  // one row group per wave and pass; a grid smaller than the groups' waves loops over them
  const int64_t stride = (int64_t)gridDim.x * 4;
  for (int64_t grp = (int64_t)blockIdx.x * 4 + wv; grp < n_grp; grp += stride) {  // uniform
    const int g = find_bucket(grp_bounds, n_grids, grp);
    const DcGridDev G = grids[g];
    const int64_t lr0 = (grp - G.sr_off) * R;
    const int nrow = (int)min<int64_t>(R, (int64_t)G.ny * G.nz - lr0);
    int ys[R], zs[R];
    {
      // lr0 < ny * nz <= 2^32 (dimensions are at most 65536): 32-bit division suffices
      const uint32_t l32 = (uint32_t)lr0, ny32 = (uint32_t)G.ny;
      int y = (int)(l32 % ny32), z = (int)(l32 / ny32);
  #pragma unroll
      for (int r = 0; r < R; ++r) {
        ys[r] = y;
        zs[r] = z;
        if (++y == G.ny) {
          y = 0;
          ++z;
        }
      }
    }
    const uint32_t lane_meta = ((uint32_t)(lane >> 4) << 12) | (uint32_t)(64 * (lane & 15));
    // rows start on 16-byte boundaries: row buffers, stores and zero fills move float4s
    const bool vec4 = (G.nx & 3) == 0 && (G.samp_off & 3) == 0 && ((uintptr_t)grad & 15) == 0;
    float carry[R];
  #pragma unroll
    for (int r = 0; r < R; ++r) carry[r] = 0.f;
    for (int s0 = 0; s0 < G.Ws; s0 += kGatherStep) {
      const int nstep = min(kGatherStep, G.Ws - s0);
      const int nsamp = min(G.nx - 64 * s0, 64 * nstep);
      uint64_t wl[R];
      uint32_t pl[R], off[R], cnt[R], base[R];
  #pragma unroll
      for (int r = 0; r < R; ++r) {
        wl[r] = 0ull;
        pl[r] = 0u;
        if (r < nrow) gather_words(G, act, wprefix, ys[r], zs[r], s0, lane, wl[r], pl[r]);
      }
      uint32_t total = 0;
  #pragma unroll
      for (int r = 0; r < R; ++r) {
        const uint32_t c = (uint32_t)__popcll(wl[r]);
        const uint32_t inc = wave_scan_incl(c, lane);
        cnt[r] = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
        base[r] = total;
        off[r] = total + inc - c;
        total += cnt[r];
      }
      auto init_row = [&](int r) {
        float4* b4 = reinterpret_cast<float4*>(buf);
        for (int x = lane; x <= 16 * nstep; x += 64) b4[x] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (lane == 0) buf[0] = carry[r];
      };
      auto store_row = [&](int r, bool summed) {
        float* out = grad + G.samp_off + ((int64_t)zs[r] * G.ny + ys[r]) * G.nx + 64 * s0;
        if (vec4) {
          float4* o4 = reinterpret_cast<float4*>(out);
          const float4* b4 = reinterpret_cast<const float4*>(buf);
          if (summed)
            for (int x = lane; x < (nsamp >> 2); x += 64) gstore(o4 + x, b4[x]);
          else
            for (int x = lane; x < (nsamp >> 2); x += 64) gstore(o4 + x, make_float4(0.f, 0.f, 0.f, 0.f));
        } else {
          if (summed)
            for (int x = lane; x < nsamp; x += 64) gstore(out + x, buf[x]);
          else
            for (int x = lane; x < nsamp; x += 64) gstore(out + x, 0.f);
        }
        if (summed) carry[r] = buf[64 * nstep];
      };
      if (total <= (uint32_t)kGatherCap) {  // uniform: every listed cell loaded at once
  #pragma unroll
        for (int r = 0; r < R; ++r)
          list_cells(cells, wl[r], pl[r], off[r], 0u, total, lane_meta | ((uint32_t)r << 14));
        wave_lds_sync();
        uint2 d[kGatherLoads];
        float2 v[kGatherLoads];
  #pragma unroll
        for (int u = 0; u < kGatherLoads; ++u) {
          d[u] = make_uint2(0u, 0u);
          v[u] = make_float2(0.f, 0.f);
          if (64u * u < total) {  // uniform
            d[u] = cells[min(64u * u + lane, total - 1u)];
            v[u] = cell_pair(contrib, d[u]);
          }
        }
  #pragma unroll
        for (int r = 0; r < R; ++r) {
          if (r >= nrow) break;  // uniform
          const bool summed = cnt[r] != 0u || carry[r] != 0.f;
          if (summed) {
            wave_lds_sync();
            init_row(r);
  #pragma unroll
            for (int u = 0; u < kGatherLoads; ++u)
              if (64u * u < total) add_cells(buf, d[u], v[u], 64u * u + lane < total, r);
            wave_lds_sync();
          }
          store_row(r, summed);
        }
      } else {  // a dense step: one row at a time, in passes of kGatherCap cells
  #pragma unroll
        for (int r = 0; r < R; ++r) {
          if (r >= nrow) break;  // uniform
          const bool summed = cnt[r] != 0u || carry[r] != 0.f;
          if (summed) {
            wave_lds_sync();
            init_row(r);
            for (uint32_t lo = base[r]; lo < base[r] + cnt[r]; lo += kGatherCap) {
              const uint32_t hi = min(base[r] + cnt[r], lo + (uint32_t)kGatherCap);
              wave_lds_sync();
              list_cells(cells, wl[r], pl[r], off[r], lo, hi, lane_meta | ((uint32_t)r << 14));
              wave_lds_sync();
              for (uint32_t k0 = 0; k0 < hi - lo; k0 += 64) {
                const uint2 d = cells[min(k0 + lane, hi - lo - 1u)];
                add_cells(buf, d, cell_pair(contrib, d), k0 + lane < hi - lo, r);
              }
            }
            wave_lds_sync();
          }
          store_row(r, summed);
        }
      }
    }
  }
}

}  // namespace

