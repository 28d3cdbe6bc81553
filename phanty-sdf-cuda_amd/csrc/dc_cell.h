// dc_cell.h — per-active-cell dual-contouring math for the CDNA4 kernels (device code).
//
// Two forward forms:
//   FP64 "plain": the reference CPU backend's arithmetic in its global coordinates, with the same
//     operation order as the CPU restatement, so vertices agree bit for bit.
//   FP32: every cell works in its own frame (min corner at the origin), the QEF is assembled
//     about the mass point with Neumaier-compensated sums, and a cell whose eigenvalues sit near
//     the 0.01 truncation threshold, whose smallest kept eigenvalue is small (poor conditioning)
//     or whose normals are degenerate is recomputed in FP64.
// The adjoint differentiates the same function at fixed topology and truncation rank.
//
// Edge and corner geometry is arithmetic (no lookup tables), so cell data stays in registers.
//
// Reference semantics (file:line in the reference tree):
//   corner order contouring.cpp:232-241; edge pairs include/Cell.h:26-30
//   hermite data Cell.cpp:162-212 (lerp_t :49-52, trilinear gradient :80-155)
//   normal averaging hermite_update.cpp:171-438 (cells per edge: Cell.cpp:26-45)
//   solve Cell.cpp:284-361, centroid :586-600
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dc_common.h"
#include "dc_internal.h"

namespace dccell {

constexpr double kSvdThreshold = 0.01;  // Cell.cpp:314
// FP32 -> FP64 fallback: an eigenvalue within kFbBand of the threshold, a kept eigenvalue below
// kFbKept (poor conditioning), or a final hermite normal that is not of unit length.
constexpr float kFbBand = 5e-4f;
constexpr float kFbKept = 0.05f;

// Corners (contouring.cpp:232-241): 0 (0,0,0) 1 (1,0,0) 2 (1,1,0) 3 (0,1,0) 4..7 the same at z=1.
__device__ __forceinline__ int cbx(int q) { return (((q & 3) + 1) >> 1) & 1; }
__device__ __forceinline__ int cby(int q) { return (q >> 1) & 1; }
__device__ __forceinline__ int cbz(int q) { return q >> 2; }
// Edges (include/Cell.h:26-30): {0,1} {1,2} {2,3} {3,0} {4,5} {5,6} {6,7} {7,4} {0,4} {1,5} {2,6}
// {3,7}.
__device__ __forceinline__ int edge_a(int e) { return e < 8 ? e : e - 8; }
__device__ __forceinline__ int edge_b(int e) {
  return e < 8 ? (((e & 3) == 3) ? e - 3 : e + 1) : e - 4;
}
// Axis of edge e (0 x, 1 y, 2 z) and the offset of its lower endpoint in the cell.
__device__ __forceinline__ int edge_axis(int e) { return e >= 8 ? 2 : (e & 1); }
__device__ __forceinline__ void edge_lower(int e, int& dx, int& dy, int& dz) {
  if (e >= 8) {
    dx = cbx(e - 8);
    dy = cby(e - 8);
    dz = 0;
  } else if ((e & 1) == 0) {  // x edges 0, 2, 4, 6
    dx = 0;
    dy = (e >> 1) & 1;
    dz = e >> 2;
  } else {  // y edges 1, 3, 5, 7
    dx = 1 - ((e >> 1) & 1);
    dy = 0;
    dz = e >> 2;
  }
}
// The 4 cells around a grid edge in the order process_edge visits them (hermite_update.cpp:
// 171-332): [base, base+dir1, base+dir2, base+dir1+dir2], base = cell whose min corner is the
// edge's lower endpoint, (dir1, dir2) = (-y,-z) for x edges, (-x,-z) for y, (-x,-y) for z
// (get_face_normals_from_edge_idx, Cell.cpp:26-45). Local edge of the grid edge in member q:
//   x: 0, 2, 4, 6     y: 3, 1, 7, 5     z: 8, 9, 11, 10
__device__ __forceinline__ void group_member(int axis, int q, int x0, int y0, int z0, int& x,
                                             int& y, int& z) {
  const int m1 = q & 1, m2 = (q >> 1) & 1;
  x = x0 - (axis == 0 ? 0 : m1);
  y = y0 - (axis == 0 ? m1 : (axis == 2 ? m2 : 0));
  z = z0 - (axis == 2 ? 0 : m2);
}
__device__ __forceinline__ int group_edge(int axis, int q) {
  if (axis == 0) return 2 * q;
  if (axis == 1) return ((q & 1) ? 1 : 3) + ((q & 2) ? 4 : 0);
  return q == 0 ? 8 : (q == 1 ? 9 : (q == 2 ? 11 : 10));
}

// dc_backend.cpp:63-65: FP32 multiply then FP32 add, widened.
__device__ __forceinline__ double coordd(float o, int i, float v) {
  return (double)__fadd_rn(o, __fmul_rn((float)i, v));
}

template <typename T>
__device__ __forceinline__ T tsqrt(T x);
template <>
__device__ __forceinline__ double tsqrt<double>(double x) {
  return sqrt(x);
}
template <>
__device__ __forceinline__ float tsqrt<float>(float x) {
  return __builtin_amdgcn_sqrtf(x);
}

// Division, reciprocal and reciprocal square root: exact IEEE operations in FP64 (the plain mode
// reproduces the CPU backend's vertices bit for bit), the hardware approximations (about 1 ulp)
// in FP32, whose results are held to a tolerance.
__device__ __forceinline__ double dv(double a, double b) { return a / b; }
__device__ __forceinline__ float dv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double rc(double b) { return 1.0 / b; }
__device__ __forceinline__ float rc(float b) { return __builtin_amdgcn_rcpf(b); }

template <typename T>
__device__ __forceinline__ T sq3(T x, T y, T z) {
  return (x * x + y * y) + z * z;
}
template <typename T>
__device__ __forceinline__ T dot3(const T a[3], const T b[3]) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
template <typename T>
__device__ __forceinline__ T tmin(T a, T b) {
  return (b < a) ? b : a;
}
template <typename T>
__device__ __forceinline__ T tmax(T a, T b) {
  return (a < b) ? b : a;
}
template <typename T>
__device__ __forceinline__ T clamp01(T t) {
  return tmax(T(0), tmin(T(1), t));
}

// S[q] += x for a run-time q without indexing the register array.
template <typename T>
__device__ __forceinline__ void add8(T* S, int q, T x) {
#pragma unroll
  for (int t = 0; t < 8; ++t) S[t] = (q == t) ? S[t] + x : S[t];
}

// fa*fb > 0 in double is exact for float inputs; FP32 uses the equivalent sign logic.
__device__ __forceinline__ bool skip_edge(double fa, double fb) { return fa * fb > 0.0; }
__device__ __forceinline__ bool skip_edge(float fa, float fb) {
  return (fa > 0.f && fb > 0.f) || (fa < 0.f && fb < 0.f);
}
__device__ __forceinline__ bool strict_change(double fa, double fb) { return fa * fb < 0.0; }
__device__ __forceinline__ bool strict_change(float fa, float fb) {
  return (fa < 0.f && fb > 0.f) || (fa > 0.f && fb < 0.f);
}

// S[] holds the corner values; s(q) for a run-time corner index picks one with a tree of three
// selects on the bits of q (indexing the register array would move it to scratch memory, and
// re-reading the sample would wait on memory again).
template <typename T>
struct Cell {
  T S[8];
  T X0, X1, Y0, Y1, Z0, Z1;
  gfloat base;
  int64_t sy, sz;
  __device__ __forceinline__ T s(int q) const {
    T v[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      v[r] = S[r];
      asm("" : "+v"(v[r]));  // a register value, so the selects stay selects
    }
    const bool b0 = q & 1, b1 = q & 2, b2 = q & 4;
    const T lo = b1 ? (b0 ? v[3] : v[2]) : (b0 ? v[1] : v[0]);
    const T hi = b1 ? (b0 ? v[7] : v[6]) : (b0 ? v[5] : v[4]);
    return b2 ? hi : lo;
  }
  __device__ __forceinline__ T cx(int q) const { return cbx(q) ? X1 : X0; }
  __device__ __forceinline__ T cy(int q) const { return cby(q) ? Y1 : Y0; }
  __device__ __forceinline__ T cz(int q) const { return cbz(q) ? Z1 : Z0; }
};

__device__ __forceinline__ gfloat sample_ptr(const DcGridDev& G, int i, int j, int k) {
  return dccommon::to_global(G.val) + ((int64_t)i + (int64_t)G.nx * ((int64_t)j + (int64_t)G.ny * k));
}

template <typename T>
__device__ __forceinline__ void load_values(const DcGridDev& G, int i, int j, int k, Cell<T>& c) {
  const gfloat b = sample_ptr(G, i, j, k);
  const int64_t sy = G.nx, sz = (int64_t)G.nx * G.ny;
  c.base = b;
  c.sy = sy;
  c.sz = sz;
  T* S = c.S;
  S[0] = (T)b[0];
  S[1] = (T)b[1];
  S[2] = (T)b[1 + sy];
  S[3] = (T)b[sy];
  S[4] = (T)b[sz];
  S[5] = (T)b[1 + sz];
  S[6] = (T)b[1 + sy + sz];
  S[7] = (T)b[sy + sz];
}

// FP64: global frame (the CPU backend's coordinates).
__device__ __forceinline__ void cell_coords(const DcGridDev& G, int i, int j, int k,
                                            Cell<double>& c) {
  c.X0 = coordd(G.ox, i, G.vx);
  c.X1 = coordd(G.ox, i + 1, G.vx);
  c.Y0 = coordd(G.oy, j, G.vy);
  c.Y1 = coordd(G.oy, j + 1, G.vy);
  c.Z0 = coordd(G.oz, k, G.vz);
  c.Z1 = coordd(G.oz, k + 1, G.vz);
}

// FP32: the cell's own frame, min corner at 0 (edge lengths from the FP32 grid coordinates).
__device__ __forceinline__ void cell_coords(const DcGridDev& G, int i, int j, int k,
                                            Cell<float>& c) {
  c.X0 = 0.f;
  c.Y0 = 0.f;
  c.Z0 = 0.f;
  c.X1 = (float)(coordd(G.ox, i + 1, G.vx) - coordd(G.ox, i, G.vx));
  c.Y1 = (float)(coordd(G.oy, j + 1, G.vy) - coordd(G.oy, j, G.vy));
  c.Z1 = (float)(coordd(G.oz, k + 1, G.vz) - coordd(G.oz, k, G.vz));
}

template <typename T>
__device__ __forceinline__ void load_cell(const DcGridDev& G, int i, int j, int k, Cell<T>& c) {
  load_values(G, i, j, k, c);
  cell_coords(G, i, j, k, c);
}

// Cell.cpp:80-155 gradientAtPointUsingTrilinear
template <typename T>
__device__ __forceinline__ void trilinear_grad(const Cell<T>& c, const T p[3], T g[3]) {
  const T S000 = c.S[0], S100 = c.S[1], S110 = c.S[2], S010 = c.S[3];
  const T S001 = c.S[4], S101 = c.S[5], S111 = c.S[6], S011 = c.S[7];
  const T dx = c.X1 - c.X0, dy = c.Y1 - c.Y0, dz = c.Z1 - c.Z0;
  const T eps = T(1e-12);
  T u = T(0.5), v = T(0.5), w = T(0.5);
  if (fabs(dx) > eps) u = dv(p[0] - c.X0, dx);
  if (fabs(dy) > eps) v = dv(p[1] - c.Y0, dy);
  if (fabs(dz) > eps) w = dv(p[2] - c.Z0, dz);
  u = clamp01(u);
  v = clamp01(v);
  w = clamp01(w);
  const T omu = T(1) - u, omv = T(1) - v, omw = T(1) - w;
  const T du = (S100 - S000) * omv * omw + (S110 - S010) * v * omw + (S101 - S001) * omv * w +
               (S111 - S011) * v * w;
  const T dvv = (S010 - S000) * omu * omw + (S110 - S100) * u * omw + (S011 - S001) * omu * w +
                (S111 - S101) * u * w;
  const T dw = (S001 - S000) * omu * omv + (S101 - S100) * u * omv + (S011 - S010) * omu * v +
               (S111 - S110) * u * v;
  g[0] = T(0);
  g[1] = T(0);
  g[2] = T(0);
  if (fabs(dx) > eps) g[0] = dv(du, dx);
  if (fabs(dy) > eps) g[1] = dv(dvv, dy);
  if (fabs(dz) > eps) g[2] = dv(dw, dz);
}

template <typename T>
__device__ __forceinline__ void normalize3(T v[3]) {
  const T z = sq3(v[0], v[1], v[2]);
  if (z > T(0)) {
    if constexpr (sizeof(T) == 8) {
      const T s = tsqrt(z);
      v[0] = v[0] / s;
      v[1] = v[1] / s;
      v[2] = v[2] / s;
    } else {
      const float r = __builtin_amdgcn_rsqf(z);
      v[0] = v[0] * r;
      v[1] = v[1] * r;
      v[2] = v[2] * r;
    }
  }
}

template <typename T>
__device__ __forceinline__ void hermite_point(const Cell<T>& c, int e, T p[3]) {
  const int a = edge_a(e), b = edge_b(e);
  const T fa = c.s(a), fb = c.s(b);
  const T t = dv(fa, fa - fb + T(1e-16));
  p[0] = (T(1) - t) * c.cx(a) + t * c.cx(b);
  p[1] = (T(1) - t) * c.cy(a) + t * c.cy(b);
  p[2] = (T(1) - t) * c.cz(a) + t * c.cz(b);
}

template <typename T>
__device__ __forceinline__ bool edge_crosses(const Cell<T>& c, int e) {
  return !skip_edge(c.s(edge_a(e)), c.s(edge_b(e)));
}

// Bit e set for each edge e that carries a hermite sample (fa*fb <= 0), from the corner values in
// registers. Loops over the set bits (ascending e, the row order of the solve) give every lane its
// own trip count, rather than all 12 edges for the whole wave.
template <typename T>
__device__ __forceinline__ uint32_t crossing_mask(const Cell<T>& c) {
  uint32_t m = 0;
#pragma unroll
  for (int e = 0; e < 12; ++e) {
    const int a = (e < 8) ? e : e - 8;
    const int b = (e < 8) ? (((e & 3) == 3) ? e - 3 : e + 1) : e - 4;
    m |= (uint32_t)(!skip_edge(c.S[a], c.S[b])) << e;
  }
  return m;
}

// Cell.cpp:180-209: the normalized trilinear normal at the hermite point of edge e (the caller
// has checked that the edge carries a sample).
template <typename T>
__device__ __forceinline__ void raw_normal(const Cell<T>& c, int e, T n[3]) {
  T p[3];
  hermite_point(c, e, p);
  trilinear_grad(c, p, n);
  normalize3(n);
}

__device__ __forceinline__ bool cell_active(const DcGridDev& G, const uint64_t* __restrict__ act,
                                            int i, int j, int k) {
  const uint64_t w = act[G.act_off + ((int64_t)k * G.cy + j) * G.W + (i >> 6)];
  return (w >> (i & 63)) & 1ull;
}

// hermite_update.cpp:171-332: is local edge e of cell (i,j,k) averaged? Returns the edge axis and
// base cell. The caller has established that e carries a hermite sample (fa*fb <= 0); a strict
// sign change makes the 4 cells active, otherwise they are looked up.
__device__ __forceinline__ bool edge_group(const DcGridDev& G, const uint64_t* __restrict__ act,
                                           int i, int j, int k, int e, bool strict, int& axis,
                                           int& x0, int& y0, int& z0) {
  int dx, dy, dz;
  edge_lower(e, dx, dy, dz);
  axis = edge_axis(e);
  x0 = i + dx;
  y0 = j + dy;
  z0 = k + dz;
  const int x1 = x0 + (axis == 0), y1 = y0 + (axis == 1), z1 = z0 + (axis == 2);
  const bool in0 = x0 > 0 && x0 < G.nx - 1 && y0 > 0 && y0 < G.ny - 1 && z0 > 0 && z0 < G.nz - 1;
  const bool in1 = x1 > 0 && x1 < G.nx - 1 && y1 > 0 && y1 < G.ny - 1 && z1 > 0 && z1 < G.nz - 1;
  if (!in0 && !in1) return false;
  if (!strict) {
    for (int q = 0; q < 4; ++q) {
      int x, y, z;
      group_member(axis, q, x0, y0, z0, x, y, z);
      if (!cell_active(G, act, x, y, z)) return false;
    }
  }
  return true;
}

// Sum of the group's nonzero raw normals in process_edge order.
template <typename T, bool kPreload = false>
__device__ __forceinline__ void group_sum(const DcGridDev& G, int axis, int x0, int y0, int z0,
                                          T acc[3]) {
  acc[0] = acc[1] = acc[2] = T(0);
  if constexpr (kPreload) {
    // all four members' corners in flight before any normal is formed (more registers: used
    // where the average is all the thread does)
    Cell<T> nc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int x, y, z;
      group_member(axis, q, x0, y0, z0, x, y, z);
      load_cell(G, x, y, z, nc[q]);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      T n[3];
      raw_normal(nc[q], group_edge(axis, q), n);
      if (sq3(n[0], n[1], n[2]) > T(0)) {
        acc[0] = acc[0] + n[0];
        acc[1] = acc[1] + n[1];
        acc[2] = acc[2] + n[2];
      }
    }
  } else {
    for (int q = 0; q < 4; ++q) {
      int x, y, z;
      group_member(axis, q, x0, y0, z0, x, y, z);
      Cell<T> nc;
      load_cell(G, x, y, z, nc);
      T n[3];
      raw_normal(nc, group_edge(axis, q), n);
      if (sq3(n[0], n[1], n[2]) > T(0)) {
        acc[0] = acc[0] + n[0];
        acc[1] = acc[1] + n[1];
        acc[2] = acc[2] + n[2];
      }
    }
  }
}

// The corner block shared by the four members of the axis-A group at base cell (x0,y0,z0):
// 2x3x3 samples (x, y, z) for an x edge, 3x2x3 for y, 3x3x2 for z, loaded row by row (9, 6 and 6
// row loads instead of 16 per-cell pair loads). All four members must lie inside the grid (true
// for a face's edge).
template <int A>
struct GroupBlock {
  static constexpr int NX = A == 0 ? 2 : 3, NY = A == 1 ? 2 : 3, NZ = A == 2 ? 2 : 3;
  float R[NZ][NY][NX];
  int xl, yl, zl;
  __device__ __forceinline__ void load(const DcGridDev& G, int x0, int y0, int z0) {
    xl = A == 0 ? x0 : x0 - 1;
    yl = A == 1 ? y0 : y0 - 1;
    zl = A == 2 ? z0 : z0 - 1;
#pragma unroll
    for (int zz = 0; zz < NZ; ++zz)
#pragma unroll
      for (int yy = 0; yy < NY; ++yy) {
        const gfloat p = sample_ptr(G, xl, yl + yy, zl + zz);
#pragma unroll
        for (int xx = 0; xx < NX; ++xx) R[zz][yy][xx] = p[xx];
      }
  }
  // member q (group_member order: x edge (x0, y0-m1, z0-m2), y edge (x0-m1, y0, z0-m2), z edge
  // (x0-m1, y0-m2, z0)) with the values load_cell gives it
  template <typename T>
  __device__ __forceinline__ void member(const DcGridDev& G, int q, Cell<T>& c) const {
    const int m1 = q & 1, m2 = (q >> 1) & 1;
    const int ox = A == 0 ? 0 : 1 - m1;
    const int oy = A == 0 ? 1 - m1 : (A == 1 ? 0 : 1 - m2);
    const int oz = A == 2 ? 0 : 1 - m2;
    T* S = c.S;
    S[0] = (T)R[oz][oy][ox];
    S[1] = (T)R[oz][oy][ox + 1];
    S[2] = (T)R[oz][oy + 1][ox + 1];
    S[3] = (T)R[oz][oy + 1][ox];
    S[4] = (T)R[oz + 1][oy][ox];
    S[5] = (T)R[oz + 1][oy][ox + 1];
    S[6] = (T)R[oz + 1][oy + 1][ox + 1];
    S[7] = (T)R[oz + 1][oy + 1][ox];
    const int x = xl + ox, y = yl + oy, z = zl + oz;
    c.base = sample_ptr(G, x, y, z);
    c.sy = G.nx;
    c.sz = (int64_t)G.nx * G.ny;
    cell_coords(G, x, y, z, c);
  }
};

// group_sum from the shared corner block: the same sum in the same order.
template <int A, typename T>
__device__ __forceinline__ void group_sum_block(const DcGridDev& G, int x0, int y0, int z0,
                                                T acc[3]) {
  GroupBlock<A> B;
  B.load(G, x0, y0, z0);
  acc[0] = acc[1] = acc[2] = T(0);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    Cell<T> c;
    B.template member<T>(G, q, c);
    T n[3];
    raw_normal(c, group_edge(A, q), n);
    if (sq3(n[0], n[1], n[2]) > T(0)) {
      acc[0] = acc[0] + n[0];
      acc[1] = acc[1] + n[1];
      acc[2] = acc[2] + n[2];
    }
  }
}

// Averaged normals precomputed once per face (k_face_normals): for face f, nrm[4f..4f+2] is the
// normalized 4-cell sum and nrm[4f+3] the length of the sum before normalization. The face id of
// the strict-crossing interior edge at base cell (x0,y0,z0) along `axis` is
// pw[w] + popc(ft[w] & below(z0)).
template <typename T>
struct FaceNormals {
  const uint64_t* ft;
  const uint32_t* pw;
  const double* nrm64;  // the table in FP64 (plain mode) or
  const float* nrm32;   // in FP32 (FP32 mode); read in either precision
  uint64_t cap;         // valid entries; 0 disables the lookup
  __device__ __forceinline__ T at(uint64_t q) const {
    return nrm64 ? (T)nrm64[q] : (T)nrm32[q];
  }
};

__device__ __forceinline__ uint64_t face_id(const DcGridDev& G, const uint64_t* __restrict__ ft,
                                            const uint32_t* __restrict__ pw, int axis, int x0,
                                            int y0, int z0) {
  const int64_t idx =
      G.ft_off + (((int64_t)axis * G.Wk + (z0 >> 6)) * G.cy + y0) * G.cx + x0;
  const uint64_t below = (1ull << (z0 & 63)) - 1ull;
  return (uint64_t)pw[idx] + (uint64_t)__popcll(ft[idx] & below);
}

// Final hermite normal of local edge e of cell (i,j,k): the face table for an averaged strict
// crossing, the 4-cell average computed here for an averaged edge with a zero endpoint, the cell's
// own normal otherwise. *face is set to the face id when the table was used, else ~0.
template <typename T>
__device__ __forceinline__ void final_normal(const DcGridDev& G, const uint64_t* __restrict__ act,
                             const FaceNormals<T>& fn, int i, int j, int k, int e,
                             const Cell<T>& own, T out[3], uint64_t* face, bool* averaged) {
  const bool strict = strict_change(own.s(edge_a(e)), own.s(edge_b(e)));
  int axis, x0, y0, z0;
  *face = ~0ull;
  *averaged = edge_group(G, act, i, j, k, e, strict, axis, x0, y0, z0);
  if (!*averaged) {
    raw_normal(own, e, out);
    return;
  }
  if (strict && fn.cap) {
    const uint64_t f = face_id(G, fn.ft, fn.pw, axis, x0, y0, z0);
    if (f < fn.cap) {
      out[0] = fn.at(4 * f + 0);
      out[1] = fn.at(4 * f + 1);
      out[2] = fn.at(4 * f + 2);
      *face = f;
      return;
    }
  }
  group_sum(G, axis, x0, y0, z0, out);
  normalize3(out);
}

// Cyclic Jacobi eigen-decomposition of a symmetric 3x3 matrix (Golub & Van Loan sec. 8.5). The
// FP64 instance has the CPU restatement's exact operation sequence.
template <typename T>
__device__ __forceinline__ T eig_tol();
template <>
__device__ __forceinline__ double eig_tol<double>() {
  return 1e-36;
}
template <>
__device__ __forceinline__ float eig_tol<float>() {
  return 1e-14f;
}
// |theta| above which theta*theta + 1 would overflow: use t = 1/(2 theta).
template <typename T>
__device__ __forceinline__ T eig_big();
template <>
__device__ __forceinline__ double eig_big<double>() {
  return 1e150;
}
template <>
__device__ __forceinline__ float eig_big<float>() {
  return 1e18f;
}

template <typename T>
__device__ __forceinline__ void sym3_eig(const T Min[3][3], T lam[3], T Q[3][3]) {
  T a[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      a[r][s] = Min[r][s];
      Q[r][s] = (r == s) ? T(1) : T(0);
    }
  for (int sweep = 0; sweep < 24; ++sweep) {
    const T off = (a[0][1] * a[0][1] + a[0][2] * a[0][2]) + a[1][2] * a[1][2];
    const T dia = (a[0][0] * a[0][0] + a[1][1] * a[1][1]) + a[2][2] * a[2][2];
    if (!(off > eig_tol<T>() * dia)) break;
#pragma unroll
    for (int pr = 0; pr < 3; ++pr) {
      const int p = (pr == 2) ? 1 : 0;
      const int q = (pr == 0) ? 1 : 2;
      const int r = 3 - p - q;
      const T apq = a[p][q];
      if (apq == T(0)) continue;
      const T theta = dv(a[q][q] - a[p][p], T(2) * apq);
      const T at = fabs(theta);
      T t;
      if (at > eig_big<T>())
        t = dv(T(0.5), theta);
      else
        t = rc(at + tsqrt(theta * theta + T(1)));
      if (theta < T(0)) t = -t;
      const T c = rc(tsqrt(t * t + T(1)));
      const T s = t * c;
      const T arp = a[r][p], arq = a[r][q];
      a[p][p] = a[p][p] - t * apq;
      a[q][q] = a[q][q] + t * apq;
      a[p][q] = T(0);
      a[q][p] = T(0);
      const T nrp = c * arp - s * arq;
      const T nrq = s * arp + c * arq;
      a[r][p] = nrp;
      a[p][r] = nrp;
      a[r][q] = nrq;
      a[q][r] = nrq;
#pragma unroll
      for (int row = 0; row < 3; ++row) {
        const T qp = Q[row][p], qq = Q[row][q];
        Q[row][p] = c * qp - s * qq;
        Q[row][q] = s * qp + c * qq;
      }
    }
  }
  lam[0] = a[0][0];
  lam[1] = a[1][1];
  lam[2] = a[2][2];
}

// ---------------------------------------------------------------------------------------------
// FP64 plain solve: Cell::update (Cell.cpp:284-361) with rows streamed in ascending edge order,
// matching the CPU restatement's sums term by term.
// ---------------------------------------------------------------------------------------------
// The FP64 solve from the rows (hermite point p, final normal nf) of the cell's crossing edges in
// ascending order; row(e, p, nf) provides edge e's row. state (unless null) receives the solve's
// state for the adjoint (kQefState doubles, the FP32 state's layout; g = r - M c = sum n n.(p-c)).
template <class Row>
__device__ __forceinline__ void solve_rows_fp64(uint32_t cross, Row&& row, double x[3],
                                                double* __restrict__ state = nullptr) {
  double M00 = 0, M01 = 0, M02 = 0, M11 = 0, M12 = 0, M22 = 0;
  double r0 = 0, r1 = 0, r2 = 0, c0 = 0, c1 = 0, c2 = 0;
  int m = 0;
  for (uint32_t em = cross; em; em &= em - 1) {
    const int e = __builtin_ctz(em);
    double p[3], nf[3];
    row(e, p, nf);
    const double b = (nf[0] * p[0] + nf[1] * p[1]) + nf[2] * p[2];
    M00 = M00 + nf[0] * nf[0];
    M01 = M01 + nf[0] * nf[1];
    M02 = M02 + nf[0] * nf[2];
    M11 = M11 + nf[1] * nf[1];
    M12 = M12 + nf[1] * nf[2];
    M22 = M22 + nf[2] * nf[2];
    r0 = r0 + nf[0] * b;
    r1 = r1 + nf[1] * b;
    r2 = r2 + nf[2] * b;
    c0 = c0 + p[0];
    c1 = c1 + p[1];
    c2 = c2 + p[2];
    ++m;
  }
  const double M[3][3] = {{M00, M01, M02}, {M01, M11, M12}, {M02, M12, M22}};
  double lam[3], Q[3][3];
  sym3_eig(M, lam, Q);
  const double dm = (double)m;
  c0 = c0 / dm;
  c1 = c1 / dm;
  c2 = c2 / dm;
  const double g0 = r0 - ((M00 * c0 + M01 * c1) + M02 * c2);
  const double g1 = r1 - ((M01 * c0 + M11 * c1) + M12 * c2);
  const double g2 = r2 - ((M02 * c0 + M12 * c1) + M22 * c2);
  if (state) {
    double2* d = reinterpret_cast<double2*>(state);
    d[0] = make_double2(c0, c1);
    d[1] = make_double2(c2, M00);
    d[2] = make_double2(M01, M02);
    d[3] = make_double2(M11, M12);
    d[4] = make_double2(M22, g0);
    d[5] = make_double2(g1, g2);
    d[6] = make_double2(lam[0], lam[1]);
    d[7] = make_double2(lam[2], Q[0][0]);
    d[8] = make_double2(Q[0][1], Q[0][2]);
    d[9] = make_double2(Q[1][0], Q[1][1]);
    d[10] = make_double2(Q[1][2], Q[2][0]);
    d[11] = make_double2(Q[2][1], Q[2][2]);
  }
  double y0 = 0.0, y1 = 0.0, y2 = 0.0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (!(fabs(lam[q]) > kSvdThreshold)) continue;
    const double qg = (Q[0][q] * g0 + Q[1][q] * g1) + Q[2][q] * g2;
    const double coef = qg / lam[q];
    y0 = y0 + coef * Q[0][q];
    y1 = y1 + coef * Q[1][q];
    y2 = y2 + coef * Q[2][q];
  }
  x[0] = c0 + y0;
  x[1] = c1 + y1;
  x[2] = c2 + y2;
}

// The row of crossing edge e: hermite point and final (unit) normal.
__device__ __forceinline__ void fp64_row(const DcGridDev& G, const uint64_t* __restrict__ act,
                                         const FaceNormals<double>& fn, int i, int j, int k,
                                         const Cell<double>& own, int e, double p[3],
                                         double nf[3]) {
  hermite_point(own, e, p);
  uint64_t face;
  bool avg;
  final_normal(G, act, fn, i, j, k, e, own, nf, &face, &avg);
  if (tsqrt(sq3(nf[0], nf[1], nf[2])) > 0) normalize3(nf);
}

__device__ __forceinline__ void solve_cell_fp64(const DcGridDev& G, const uint64_t* __restrict__ act,
                                const FaceNormals<double>& fn, int i, int j, int k, double x[3]) {
  Cell<double> own;
  load_cell(G, i, j, k, own);
  solve_rows_fp64(crossing_mask(own),
                  [&](int e, double p[3], double nf[3]) { fp64_row(G, act, fn, i, j, k, own, e, p, nf); },
                  x);
}

// Neumaier compensated sum.
struct NSum {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float x) {
    const float t = s + x;
    if (fabsf(s) >= fabsf(x))
      c += (s - t) + x;
    else
      c += (x - t) + s;
    s = t;
  }
  __device__ __forceinline__ float val() const { return s + c; }
};

template <typename T>
struct Qef {
  T c[3];
  T M[3][3];
  T g[3];
  int m;
};

// Kinds of a cell's rows, recorded for the adjoint: a face id (the row's normal is that face's
// average), or one of these.
constexpr uint32_t kRowOwn = 0xffffffffu;       // not averaged: the cell's own normal
constexpr uint32_t kRowEdgeSlot = 0xfffffffeu;  // averaged, without a face
// Threads of the FP32 solve's blocks: stride of a thread's kinds column in LDS.
constexpr int kKindStride = 256;

// Mass-point-centred assembly in the cell frame: centroid c, M = sum n n^T, g = sum n (n.(p-c)),
// Neumaier-compensated. Returns false when a final normal is degenerate (then solve in FP64).
// kinds (or null): each crossing row's kind for the adjoint (the face id, kRowEdgeSlot or kRowOwn)
// at kinds[e * kKindStride] (the caller's LDS column).
__device__ __forceinline__ bool assemble_fp32(const DcGridDev& G, const uint64_t* __restrict__ act,
                              const FaceNormals<float>& fn, int i, int j, int k,
                              const Cell<float>& own, Qef<float>& Q,
                              uint32_t* __restrict__ kinds) {
  NSum s0, s1, s2;
  int m = 0;
  for (uint32_t em = crossing_mask(own); em; em &= em - 1) {
    const int e = __builtin_ctz(em);
    float p[3];
    hermite_point(own, e, p);
    s0.add(p[0]);
    s1.add(p[1]);
    s2.add(p[2]);
    ++m;
  }
  const float fm = (float)m;
  Q.m = m;
  Q.c[0] = dv(s0.val(), fm);
  Q.c[1] = dv(s1.val(), fm);
  Q.c[2] = dv(s2.val(), fm);
  NSum M00, M01, M02, M11, M12, M22, g0, g1, g2;
  bool ok = true;
  for (uint32_t em = crossing_mask(own); em; em &= em - 1) {
    const int e = __builtin_ctz(em);
    float p[3];
    hermite_point(own, e, p);
    float nf[3];
    uint64_t face;
    bool avg;
    final_normal(G, act, fn, i, j, k, e, own, nf, &face, &avg);
    if (kinds) kinds[e * kKindStride] = !avg ? kRowOwn : (face != ~0ull ? (uint32_t)face : kRowEdgeSlot);
    if (!(sq3(nf[0], nf[1], nf[2]) > 0.5f)) ok = false;  // zero or degenerate normal
    normalize3(nf);
    const float q[3] = {p[0] - Q.c[0], p[1] - Q.c[1], p[2] - Q.c[2]};
    const float d = dot3(nf, q);
    M00.add(nf[0] * nf[0]);
    M01.add(nf[0] * nf[1]);
    M02.add(nf[0] * nf[2]);
    M11.add(nf[1] * nf[1]);
    M12.add(nf[1] * nf[2]);
    M22.add(nf[2] * nf[2]);
    g0.add(nf[0] * d);
    g1.add(nf[1] * d);
    g2.add(nf[2] * d);
  }
  Q.M[0][0] = M00.val();
  Q.M[0][1] = Q.M[1][0] = M01.val();
  Q.M[0][2] = Q.M[2][0] = M02.val();
  Q.M[1][1] = M11.val();
  Q.M[1][2] = Q.M[2][1] = M12.val();
  Q.M[2][2] = M22.val();
  Q.g[0] = g0.val();
  Q.g[1] = g1.val();
  Q.g[2] = g2.val();
  return ok;
}

// The FP32 solve's state, kept for the adjoint (kQefState floats per vertex, six float4s): mass
// point c, M (upper triangle), g, the eigenvalues and the eigenvectors (V[r][q], row-major).
constexpr int kQefState = 24;

__device__ __forceinline__ void store_state(float* __restrict__ dst, const Qef<float>& Q,
                                            const float lam[3], const float V[3][3]) {
  float4* d = reinterpret_cast<float4*>(dst);
  d[0] = make_float4(Q.c[0], Q.c[1], Q.c[2], Q.M[0][0]);
  d[1] = make_float4(Q.M[0][1], Q.M[0][2], Q.M[1][1], Q.M[1][2]);
  d[2] = make_float4(Q.M[2][2], Q.g[0], Q.g[1], Q.g[2]);
  d[3] = make_float4(lam[0], lam[1], lam[2], V[0][0]);
  d[4] = make_float4(V[0][1], V[0][2], V[1][0], V[1][1]);
  d[5] = make_float4(V[1][2], V[2][0], V[2][1], V[2][2]);
}

__device__ __forceinline__ void load_state(const float* __restrict__ src, Qef<float>& Q,
                                           float lam[3], float V[3][3]) {
  const float4* s = reinterpret_cast<const float4*>(src);
  const float4 a = s[0], b = s[1], c = s[2], d = s[3], e = s[4], f = s[5];
  Q.c[0] = a.x;
  Q.c[1] = a.y;
  Q.c[2] = a.z;
  Q.M[0][0] = a.w;
  Q.M[0][1] = Q.M[1][0] = b.x;
  Q.M[0][2] = Q.M[2][0] = b.y;
  Q.M[1][1] = b.z;
  Q.M[1][2] = Q.M[2][1] = b.w;
  Q.M[2][2] = c.x;
  Q.g[0] = c.y;
  Q.g[1] = c.z;
  Q.g[2] = c.w;
  lam[0] = d.x;
  lam[1] = d.y;
  lam[2] = d.z;
  V[0][0] = d.w;
  V[0][1] = e.x;
  V[0][2] = e.y;
  V[1][0] = e.z;
  V[1][1] = e.w;
  V[1][2] = f.x;
  V[2][0] = f.y;
  V[2][1] = f.z;
  V[2][2] = f.w;
}

// The FP64 fallback solve's state (solve_rows_fp64), same layout.
__device__ __forceinline__ void load_state(const double* __restrict__ src, Qef<double>& Q,
                                           double lam[3], double V[3][3]) {
  const double2* s = reinterpret_cast<const double2*>(src);
  double t[24];
#pragma unroll
  for (int q = 0; q < 12; ++q) {
    const double2 u = s[q];
    t[2 * q] = u.x;
    t[2 * q + 1] = u.y;
  }
  Q.c[0] = t[0];
  Q.c[1] = t[1];
  Q.c[2] = t[2];
  Q.M[0][0] = t[3];
  Q.M[0][1] = Q.M[1][0] = t[4];
  Q.M[0][2] = Q.M[2][0] = t[5];
  Q.M[1][1] = t[6];
  Q.M[1][2] = Q.M[2][1] = t[7];
  Q.M[2][2] = t[8];
  Q.g[0] = t[9];
  Q.g[1] = t[10];
  Q.g[2] = t[11];
  lam[0] = t[12];
  lam[1] = t[13];
  lam[2] = t[14];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) V[r][c] = t[15 + 3 * r + c];
}

// FP32 solve in the cell frame. Returns false when the cell must be solved in FP64; otherwise
// stores the solve's state for the adjoint at `state` (kQefState floats) unless it is null. The
// row kinds go to `kinds` (unless null) either way.
__device__ __forceinline__ bool solve_cell_fp32(const DcGridDev& G, const uint64_t* __restrict__ act,
                                const FaceNormals<float>& fn, int i, int j, int k, float xl[3],
                                float* __restrict__ state, uint32_t* __restrict__ kinds) {
  Cell<float> own;
  load_cell(G, i, j, k, own);
  Qef<float> Q;
  if (!assemble_fp32(G, act, fn, i, j, k, own, Q, kinds)) return false;
  float lam[3], V[3][3];
  sym3_eig(Q.M, lam, V);
  float y[3] = {0.f, 0.f, 0.f};
  bool ok = true;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const float l = lam[q];
    if (fabsf(fabsf(l) - (float)kSvdThreshold) < kFbBand) ok = false;
    if (!(fabsf(l) > (float)kSvdThreshold)) continue;
    if (l < kFbKept) ok = false;
    const float coef = dv(V[0][q] * Q.g[0] + V[1][q] * Q.g[1] + V[2][q] * Q.g[2], l);
    y[0] += coef * V[0][q];
    y[1] += coef * V[1][q];
    y[2] += coef * V[2][q];
  }
  xl[0] = Q.c[0] + y[0];
  xl[1] = Q.c[1] + y[1];
  xl[2] = Q.c[2] + y[2];
  if (ok && state) store_state(state, Q, lam, V);
  return ok;
}

// ---------------------------------------------------------------------------------------------
// Adjoint
// ---------------------------------------------------------------------------------------------
// VJP of the trilinear gradient with respect to the 8 corner values and the point.
template <typename T>
__device__ __forceinline__ void trilinear_vjp(const Cell<T>& c, const T p[3], const T Gb[3], T Sb[8], T pb[3]) {
  const T S000 = c.S[0], S100 = c.S[1], S110 = c.S[2], S010 = c.S[3];
  const T S001 = c.S[4], S101 = c.S[5], S111 = c.S[6], S011 = c.S[7];
  const T dx = c.X1 - c.X0, dy = c.Y1 - c.Y0, dz = c.Z1 - c.Z0;
  const T eps = T(1e-12);
  const bool hx = fabs(dx) > eps, hy = fabs(dy) > eps, hz = fabs(dz) > eps;
  T u = T(0.5), v = T(0.5), w = T(0.5);
  if (hx) u = dv(p[0] - c.X0, dx);
  if (hy) v = dv(p[1] - c.Y0, dy);
  if (hz) w = dv(p[2] - c.Z0, dz);
  const bool iu = hx && u >= T(0) && u <= T(1);
  const bool iv = hy && v >= T(0) && v <= T(1);
  const bool iw = hz && w >= T(0) && w <= T(1);
  u = clamp01(u);
  v = clamp01(v);
  w = clamp01(w);
  const T omu = T(1) - u, omv = T(1) - v, omw = T(1) - w;
  const T a = hx ? dv(Gb[0], dx) : T(0);
  const T bb = hy ? dv(Gb[1], dy) : T(0);
  const T cc = hz ? dv(Gb[2], dz) : T(0);
  Sb[1] += a * omv * omw; Sb[0] -= a * omv * omw;
  Sb[2] += a * v * omw;   Sb[3] -= a * v * omw;
  Sb[5] += a * omv * w;   Sb[4] -= a * omv * w;
  Sb[6] += a * v * w;     Sb[7] -= a * v * w;
  Sb[3] += bb * omu * omw; Sb[0] -= bb * omu * omw;
  Sb[2] += bb * u * omw;   Sb[1] -= bb * u * omw;
  Sb[7] += bb * omu * w;   Sb[4] -= bb * omu * w;
  Sb[6] += bb * u * w;     Sb[5] -= bb * u * w;
  Sb[4] += cc * omu * omv; Sb[0] -= cc * omu * omv;
  Sb[5] += cc * u * omv;   Sb[1] -= cc * u * omv;
  Sb[7] += cc * omu * v;   Sb[3] -= cc * omu * v;
  Sb[6] += cc * u * v;     Sb[2] -= cc * u * v;
  const T ddu_dv = ((S110 - S010) - (S100 - S000)) * omw + ((S111 - S011) - (S101 - S001)) * w;
  const T ddu_dw = ((S101 - S001) - (S100 - S000)) * omv + ((S111 - S011) - (S110 - S010)) * v;
  const T ddv_du = ((S110 - S100) - (S010 - S000)) * omw + ((S111 - S101) - (S011 - S001)) * w;
  const T ddv_dw = ((S011 - S001) - (S010 - S000)) * omu + ((S111 - S101) - (S110 - S100)) * u;
  const T ddw_du = ((S101 - S100) - (S001 - S000)) * omv + ((S111 - S110) - (S011 - S010)) * v;
  const T ddw_dv = ((S011 - S010) - (S001 - S000)) * omu + ((S111 - S110) - (S101 - S100)) * u;
  const T ub = bb * ddv_du + cc * ddw_du;
  const T vb = a * ddu_dv + cc * ddw_dv;
  const T wb = a * ddu_dw + bb * ddv_dw;
  if (iu) pb[0] += dv(ub, dx);
  if (iv) pb[1] += dv(vb, dy);
  if (iw) pb[2] += dv(wb, dz);
}

// VJP of n = v/|v| (identity when |v| = 0, as Eigen's normalize leaves v unchanged).
template <typename T>
__device__ __forceinline__ void normalize_vjp(const T v[3], const T nb[3], T vb[3]) {
  const T z = sq3(v[0], v[1], v[2]);
  if (!(z > T(0))) {
    vb[0] = nb[0];
    vb[1] = nb[1];
    vb[2] = nb[2];
    return;
  }
  const T rs = rc(tsqrt(z));
  const T n[3] = {v[0] * rs, v[1] * rs, v[2] * rs};
  const T d = dot3(n, nb);
  vb[0] = (nb[0] - n[0] * d) * rs;
  vb[1] = (nb[1] - n[1] * d) * rs;
  vb[2] = (nb[2] - n[2] * d) * rs;
}

// Adjoint of t = fa/(fa-fb+1e-16) for a point adjoint pb (p = (1-t) pa + t pb).
template <typename T>
__device__ __forceinline__ void lerp_vjp(const Cell<T>& c, int e, const T pb[3], T Sb[8]) {
  const int a = edge_a(e), b = edge_b(e);
  const T tb = (pb[0] * (c.cx(b) - c.cx(a)) + pb[1] * (c.cy(b) - c.cy(a))) +
               pb[2] * (c.cz(b) - c.cz(a));
  const T fa = c.s(a), fb = c.s(b);
  const T D = fa - fb + T(1e-16);
  const T D2 = D * D;
  const T rD2 = rc(D2);
  add8(Sb, a, tb * (T(1e-16) - fb) * rD2);
  add8(Sb, b, tb * fa * rD2);
}

// Hermite sample (cell, e): adjoint nb of its normalized normal plus point adjoint pb_in -> the
// cell's 8 corner values.
template <typename T>
__device__ __forceinline__ void hermite_vjp(const Cell<T>& c, int e, const T nb[3], const T pb_in[3], T Sb[8]) {
  T p[3], Gr[3];
  hermite_point(c, e, p);
  trilinear_grad(c, p, Gr);
  T Gb[3];
  normalize_vjp(Gr, nb, Gb);
  T pb[3] = {pb_in[0], pb_in[1], pb_in[2]};
  trilinear_vjp(c, p, Gb, Sb, pb);
  lerp_vjp(c, e, pb, Sb);
}

// Assembly in T precision for the adjoint (cell frame for float, global frame for double).
template <typename T>
__device__ __forceinline__ void assemble_any(const DcGridDev& G, const uint64_t* __restrict__ act,
                             const FaceNormals<T>& fn, int i, int j, int k, const Cell<T>& own,
                             Qef<T>& Q) {
  T c[3] = {T(0), T(0), T(0)};
  int m = 0;
  for (uint32_t em = crossing_mask(own); em; em &= em - 1) {
    const int e = __builtin_ctz(em);
    T p[3];
    hermite_point(own, e, p);
    c[0] += p[0];
    c[1] += p[1];
    c[2] += p[2];
    ++m;
  }
  Q.m = m;
#pragma unroll
  for (int d = 0; d < 3; ++d) Q.c[d] = dv(c[d], T(m));
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    Q.g[r] = T(0);
#pragma unroll
    for (int s = 0; s < 3; ++s) Q.M[r][s] = T(0);
  }
  for (uint32_t em = crossing_mask(own); em; em &= em - 1) {
    const int e = __builtin_ctz(em);
    T p[3], nf[3];
    hermite_point(own, e, p);
    uint64_t face;
    bool avg;
    final_normal(G, act, fn, i, j, k, e, own, nf, &face, &avg);
    normalize3(nf);
    const T q[3] = {p[0] - Q.c[0], p[1] - Q.c[1], p[2] - Q.c[2]};
    const T d = dot3(nf, q);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      Q.g[r] += nf[r] * d;
#pragma unroll
      for (int s = 0; s < 3; ++s) Q.M[r][s] += nf[r] * nf[s];
    }
  }
}

// Global vertex id of active cell (i,j,k): rank in the batch's vertex order.
__device__ __forceinline__ uint64_t global_vertex(const DcGridDev& G,
                                                  const uint64_t* __restrict__ act,
                                                  const uint32_t* __restrict__ wprefix, int i,
                                                  int j, int k) {
  const int64_t w = G.act_off + ((int64_t)k * G.cy + j) * G.W + (i >> 6);
  const uint64_t below = (1ull << (i & 63)) - 1ull;
  return (uint64_t)wprefix[w] + (uint64_t)__popcll(act[w] & below);
}

// Member index of cell (i,j,k) in the group of grid edge (axis, x0,y0,z0): bit 0 = moved along
// dir1, bit 1 = along dir2.
__device__ __forceinline__ int member_of(int axis, int i, int j, int k, int x0, int y0, int z0) {
  const int d1moved = (axis == 0) ? (j != y0) : (i != x0);
  const int d2moved = (axis == 2) ? (j != y0) : (k != z0);
  return d1moved + 2 * d2moved;
}

// Adjoint phase 1 for vertex v = active cell (i,j,k). For each row: the adjoint u-bar of its
// final hermite normal and p-bar of its hermite point. A row whose normal is a 4-cell average
// stores u-bar for phase 2 (in the face's slot for this member, ubar[4*face + member], or, for
// an averaged edge without a face, in the cell's per-edge slot ubar_e[12*v + e]); the point
// adjoints and the other rows' normals go to the cell's own 8 corners, Sb.
// A row's kind: the face id, kRowEdgeSlot (averaged, no face) or kRowOwn (not averaged).
// kGiven (FP32 mode): state is the forward's solve state of this cell (FP32, or FP64 for a cell
// the forward solved in FP64) and rows[12v + e] the kinds the forward recorded; the first kPre
// face rows' normals are loaded together before the first row is processed. Otherwise the cell is
// assembled and solved again here, and the kinds go to kinds_out[e * kKindStride] (the caller's
// LDS column) unless it is null.

// Entry e of a cell's 12 row kinds held as three uint4.
__device__ __forceinline__ uint32_t row_kind(const uint4 r[3], int e) {
  const uint4 q = e < 4 ? r[0] : (e < 8 ? r[1] : r[2]);
  const int c = e & 3;
  return c == 0 ? q.x : (c == 1 ? q.y : (c == 2 ? q.z : q.w));
}

// Crossing rows with a face (fmask) and with an edge slot (smask), from the 12 kinds.
// Rows of the crossing mask with a face (fmask) and with an edge slot (smask). A kind is a face id
// only below nf (the face count): a kind outside that range (never written by a consistent
// forward) names no slot and the row is taken as the cell's own normal.
__device__ __forceinline__ void kind_masks(const uint4 kinds[3], uint32_t cross, uint64_t nf,
                                           uint32_t& fmask, uint32_t& smask) {
  fmask = smask = 0;
#pragma unroll
  for (int e = 0; e < 12; ++e) {
    const uint32_t kind = row_kind(kinds, e);
    fmask |= (uint32_t)((uint64_t)kind < nf) << e;
    smask |= (uint32_t)(kind == kRowEdgeSlot) << e;
  }
  fmask &= cross;
  smask &= cross;
}

// The solve's adjoint terms shared by a cell's rows, for a vertex adjoint xb: gb = P xb (P the
// thresholded pseudo-inverse), M-bar (Daleckii-Krein) and c-bar / m.
template <typename T>
struct RowsAdjoint {
  T c[3], gb[3], Mb[3][3], cbm[3];
};

template <typename T>
__device__ __forceinline__ void rows_adjoint(const Qef<T>& Q, const T lam[3], const T V[3][3],
                                             const float xb_in[3], RowsAdjoint<T>& R) {
  const T xb[3] = {(T)xb_in[0], (T)xb_in[1], (T)xb_in[2]};
  bool kept[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    R.c[d] = Q.c[d];
    R.gb[d] = T(0);
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    kept[q] = fabs(lam[q]) > T(kSvdThreshold);
    if (!kept[q]) continue;
    const T qx = (V[0][q] * xb[0] + V[1][q] * xb[1]) + V[2][q] * xb[2];
#pragma unroll
    for (int d = 0; d < 3; ++d) R.gb[d] += V[d][q] * dv(qx, lam[q]);
  }
  // Daleckii-Krein: M-bar = V (F o (V^T sym(xb g^T) V)) V^T, F_ab = (f(l_a)-f(l_b))/(l_a-l_b),
  // f(l) = 1/l above the threshold, 0 below.
  T S[3][3];
#pragma unroll
  for (int u = 0; u < 3; ++u)
#pragma unroll
    for (int w = 0; w < 3; ++w) S[u][w] = T(0.5) * (xb[u] * Q.g[w] + Q.g[u] * xb[w]);
  T B[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      T s = T(0);
#pragma unroll
      for (int u = 0; u < 3; ++u)
#pragma unroll
        for (int w = 0; w < 3; ++w) s += V[u][a] * S[u][w] * V[w][b];
      T F;
      if (kept[a] && kept[b])
        F = -rc(lam[a] * lam[b]);
      else if (!kept[a] && !kept[b])
        F = T(0);
      else
        F = dv((kept[a] ? rc(lam[a]) : T(0)) - (kept[b] ? rc(lam[b]) : T(0)), lam[a] - lam[b]);
      B[a][b] = F * s;
    }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      T s = T(0);
#pragma unroll
      for (int u = 0; u < 3; ++u)
#pragma unroll
        for (int w = 0; w < 3; ++w) s += V[a][u] * B[u][w] * V[b][w];
      R.Mb[a][b] = s;
    }
  // x = c + P g with g = sum n (n.(p-c)): c-bar = xb - M gb, shared by the rows' points.
#pragma unroll
  for (int a = 0; a < 3; ++a)
    R.cbm[a] =
        dv(xb[a] - ((Q.M[a][0] * R.gb[0] + Q.M[a][1] * R.gb[1]) + Q.M[a][2] * R.gb[2]), T(Q.m));
}

// One row of the adjoint: edge e with final normal nf. slot: where an averaged row's u-bar goes
// (null: not averaged; its normal's adjoint goes to the cell's corners).
template <typename T>
__device__ __forceinline__ void row_adjoint(const Cell<T>& own, const RowsAdjoint<T>& R, int e,
                                            const T nf[3], float* __restrict__ slot, T Sb[8]) {
  T p[3];
  hermite_point(own, e, p);
  T n[3] = {nf[0], nf[1], nf[2]};
  normalize3(n);
  const T q[3] = {p[0] - R.c[0], p[1] - R.c[1], p[2] - R.c[2]};
  const T nq = dot3(n, q), ng = dot3(n, R.gb);
  T nb[3], pb[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const T Mn =
        (T(2) * R.Mb[d][0] * n[0] + T(2) * R.Mb[d][1] * n[1]) + T(2) * R.Mb[d][2] * n[2];
    nb[d] = Mn + R.gb[d] * nq + q[d] * ng;
    pb[d] = n[d] * ng + R.cbm[d];
  }
  T ub[3];
  normalize_vjp(nf, nb, ub);  // the solve re-normalizes each row
  if (!slot) {
    hermite_vjp(own, e, ub, pb, Sb);
    return;
  }
  lerp_vjp(own, e, pb, Sb);  // the point enters the solve directly
  slot[0] = (float)ub[0];
  slot[1] = (float)ub[1];
  slot[2] = (float)ub[2];
}

// Where an averaged row's u-bar goes: the face's slot for this member of its group.
__device__ __forceinline__ float* face_slot(const DcGridDev& G, const uint64_t* __restrict__ act,
                                            int i, int j, int k, int e, uint32_t face,
                                            float* __restrict__ ubar) {
  int axis, x0, y0, z0;  // the group's geometry (the row is averaged)
  edge_group(G, act, i, j, k, e, true, axis, x0, y0, z0);
  return ubar + 3 * (4 * (uint64_t)face + member_of(axis, i, j, k, x0, y0, z0));
}

// Final normal and u-bar slot of a row whose kind is not a face (kRowEdgeSlot: the 4-cell average
// formed here; kRowOwn: the cell's own normal, no slot).
template <typename T>
__device__ __forceinline__ float* row_nonface(const DcGridDev& G, const uint64_t* __restrict__ act,
                                              int i, int j, int k, const Cell<T>& own, int e,
                                              uint32_t kind, uint64_t v, float* __restrict__ ubar_e,
                                              T nf[3]) {
  if (kind == kRowEdgeSlot) {
    int axis, x0, y0, z0;
    edge_group(G, act, i, j, k, e, true, axis, x0, y0, z0);
    group_sum(G, axis, x0, y0, z0, nf);
    normalize3(nf);
    return ubar_e + 3 * (12 * v + e);
  }
  raw_normal(own, e, nf);
  return nullptr;
}

template <typename T, bool kGiven>
__device__ __forceinline__ void backward_rows(const DcGridDev& G, const uint64_t* __restrict__ act,
                              const FaceNormals<T>& fn, int i, int j, int k, uint64_t v,
                              const float xb_in[3], const T* __restrict__ state,
                              float* __restrict__ ubar, float* __restrict__ ubar_e,
                              const uint32_t* __restrict__ rows, uint32_t* __restrict__ kinds_out,
                              T Sb[8]) {
  uint4 kinds[3];
  if constexpr (kGiven) {
    const uint4* r4 = reinterpret_cast<const uint4*>(rows + 12 * v);
    kinds[0] = r4[0];
    kinds[1] = r4[1];
    kinds[2] = r4[2];
  }
  Cell<T> own;
  load_cell(G, i, j, k, own);
  Qef<T> Q;
  T lam[3], V[3][3];
  if constexpr (kGiven) {
    load_state(state, Q, lam, V);
    Q.m = __popc(crossing_mask(own));
  } else {
    assemble_any(G, act, fn, i, j, k, own, Q);
    sym3_eig(Q.M, lam, V);
  }
  RowsAdjoint<T> R;
  rows_adjoint(Q, lam, V, xb_in, R);
#pragma unroll
  for (int q = 0; q < 8; ++q) Sb[q] = T(0);
  const uint32_t cross = crossing_mask(own);
  if constexpr (kGiven) {
    constexpr int kPre = 3;
    uint32_t fmask, smask;
    kind_masks(kinds, cross, fn.cap, fmask, smask);
    T pre[kPre][3];
    uint32_t pf[kPre];
    {
      uint32_t pm = fmask;
#pragma unroll
      for (int q = 0; q < kPre; ++q) {
        pre[q][0] = pre[q][1] = pre[q][2] = T(0);
        pf[q] = 0;
        if (pm) {
          pf[q] = row_kind(kinds, __builtin_ctz(pm));
          pre[q][0] = fn.at(4 * (uint64_t)pf[q] + 0);
          pre[q][1] = fn.at(4 * (uint64_t)pf[q] + 1);
          pre[q][2] = fn.at(4 * (uint64_t)pf[q] + 2);
          pm &= pm - 1;
        }
      }
    }
    int q = 0;  // face rows seen
    for (uint32_t em = cross; em; em &= em - 1) {
      const int e = __builtin_ctz(em);
      T nf[3];
      float* slot;
      if ((fmask >> e) & 1u) {
        uint32_t face;
        if (q < kPre) {
          face = q == 0 ? pf[0] : (q == 1 ? pf[1] : pf[2]);
#pragma unroll
          for (int d = 0; d < 3; ++d)
            nf[d] = q == 0 ? pre[0][d] : (q == 1 ? pre[1][d] : pre[2][d]);
        } else {
          face = rows[12 * v + e];
          nf[0] = fn.at(4 * (uint64_t)face + 0);
          nf[1] = fn.at(4 * (uint64_t)face + 1);
          nf[2] = fn.at(4 * (uint64_t)face + 2);
        }
        ++q;
        slot = face_slot(G, act, i, j, k, e, face, ubar);
      } else {
        slot = row_nonface(G, act, i, j, k, own, e, ((smask >> e) & 1u) ? kRowEdgeSlot : kRowOwn,
                           v, ubar_e, nf);
      }
      row_adjoint(own, R, e, nf, slot, Sb);
    }
  } else {
    for (uint32_t em = cross; em; em &= em - 1) {
      const int e = __builtin_ctz(em);
      T nf[3];
      uint64_t face;
      bool avg;
      final_normal(G, act, fn, i, j, k, e, own, nf, &face, &avg);
      if (kinds_out)
        kinds_out[e * kKindStride] =
            !avg ? kRowOwn : (face != ~0ull ? (uint32_t)face : kRowEdgeSlot);
      float* slot = nullptr;
      if (avg)
        slot = face != ~0ull ? face_slot(G, act, i, j, k, e, (uint32_t)face, ubar)
                             : ubar_e + 3 * (12 * v + e);
      row_adjoint(own, R, e, nf, slot, Sb);
    }
  }
}

// The same for one row e of a cell whose kinds the forward recorded (kind: the row's kind), with
// the cell's RowsAdjoint: for one lane of a cell's lane group.
template <typename T>
__device__ __forceinline__ void backward_row_given(const DcGridDev& G,
                                                   const uint64_t* __restrict__ act,
                                                   const FaceNormals<T>& fn, int i, int j, int k,
                                                   uint64_t v, const Cell<T>& own,
                                                   const RowsAdjoint<T>& R, int e, uint32_t kind,
                                                   float* __restrict__ ubar,
                                                   float* __restrict__ ubar_e, T Sb[8]) {
  T nf[3];
  float* slot;
  if ((uint64_t)kind < fn.cap) {
    nf[0] = fn.at(4 * (uint64_t)kind + 0);
    nf[1] = fn.at(4 * (uint64_t)kind + 1);
    nf[2] = fn.at(4 * (uint64_t)kind + 2);
    slot = face_slot(G, act, i, j, k, e, kind, ubar);
  } else {  // an edge slot, or the own normal (also for a kind past the faces)
    slot = row_nonface(G, act, i, j, k, own, e, kind == kRowEdgeSlot ? kRowEdgeSlot : kRowOwn, v,
                       ubar_e, nf);
  }
  row_adjoint(own, R, e, nf, slot, Sb);
}

// Between the phases, per face f: the adjoint of the face's averaged (normalized) normal taken
// back to the 4-cell sum, from the members' u-bar summed in member order, computed in FP64 and
// written as S (double: plain mode, float: FP32 mode) over the face's first slot floats.
template <typename S>
__device__ __forceinline__ void face_adjoint(const FaceNormals<double>& fn, uint64_t f,
                                             float* __restrict__ ubar) {
  const float4* u = reinterpret_cast<const float4*>(ubar + 12 * f);
  const float4 a = u[0], b = u[1], c = u[2];
  // slots m = 0..3: (a.xyz), (a.w, b.xy), (b.zw, c.x), (c.yzw)
  const double us[3] = {(((double)a.x + (double)a.w) + (double)b.z) + (double)c.y,
                        (((double)a.y + (double)b.x) + (double)b.w) + (double)c.z,
                        (((double)a.z + (double)b.y) + (double)c.x) + (double)c.w};
  const double len = fn.at(4 * f + 3);
  const double acc[3] = {fn.at(4 * f + 0) * len, fn.at(4 * f + 1) * len, fn.at(4 * f + 2) * len};
  double accb[3];
  normalize_vjp(acc, us, accb);
  S* o = reinterpret_cast<S*>(ubar + 12 * f);
  o[0] = (S)accb[0];
  o[1] = (S)accb[1];
  o[2] = (S)accb[2];
}

// Adjoint of the 4-cell normal sum of averaged edge e of cell (i,j,k) that has no face: the
// u-bar of the group's 4 cells (their per-edge slots) summed and taken through the average's
// normalization.
template <typename T>
__device__ __forceinline__ void edge_slot_adjoint(const DcGridDev& G,
                                                  const uint64_t* __restrict__ act,
                                                  const uint32_t* __restrict__ wprefix, int i,
                                                  int j, int k, int e,
                                                  const float* __restrict__ ubar_e, T accb[3]) {
  int axis, x0, y0, z0;  // the group's geometry only (phase 1 found it averaged)
  edge_group(G, act, i, j, k, e, true, axis, x0, y0, z0);
  T us[3] = {T(0), T(0), T(0)};
  for (int m = 0; m < 4; ++m) {
    int x, y, z;
    group_member(axis, m, x0, y0, z0, x, y, z);
    const uint64_t vm = global_vertex(G, act, wprefix, x, y, z);
    const float* slot = ubar_e + 3 * (12 * vm + group_edge(axis, m));
    us[0] += (T)slot[0];
    us[1] += (T)slot[1];
    us[2] += (T)slot[2];
  }
  T acc[3];
  group_sum(G, axis, x0, y0, z0, acc);
  normalize_vjp(acc, us, accb);
}

// Adjoint phase 2 for one active cell: for each of its averaged rows, the adjoint of the group's
// 4-cell normal sum (per face: computed once by face_adjoint, stored as S; for an averaged edge
// without a face: the u-bar of the group's 4 cells summed and taken through the average's
// normalization here), then through this cell's own raw normal (its share of the average) into
// Sb. rows: the cell's 12 row kinds written by phase 1. The first kPre face rows' adjoints are
// loaded together before the first row is processed.
template <typename T, typename S>
__device__ __forceinline__ void backward_faces(const DcGridDev& G, const uint64_t* __restrict__ act,
                               const uint32_t* __restrict__ wprefix, int i, int j, int k,
                               const uint32_t* __restrict__ rows, const float* __restrict__ ubar,
                               const float* __restrict__ ubar_e, uint64_t nf, T Sb[8]) {
  constexpr int kPre = 3;
  const uint4* r4 = reinterpret_cast<const uint4*>(rows);
  const uint4 kinds[3] = {r4[0], r4[1], r4[2]};
  Cell<T> own;
  load_cell(G, i, j, k, own);
  uint32_t fmask, smask;  // crossing rows with a face, with an edge slot
  kind_masks(kinds, crossing_mask(own), nf, fmask, smask);
  T pre[kPre][3];
  {
    uint32_t pm = fmask;
#pragma unroll
    for (int p = 0; p < kPre; ++p) {
      pre[p][0] = pre[p][1] = pre[p][2] = T(0);
      if (pm) {
        const S* a =
            reinterpret_cast<const S*>(ubar + 12 * (uint64_t)row_kind(kinds, __builtin_ctz(pm)));
        pre[p][0] = (T)a[0];
        pre[p][1] = (T)a[1];
        pre[p][2] = (T)a[2];
        pm &= pm - 1;
      }
    }
  }
  const T zero3[3] = {T(0), T(0), T(0)};
  int p = 0;  // face rows seen
  for (uint32_t em = fmask | smask; em; em &= em - 1) {
    const int e = __builtin_ctz(em);
    const bool has_face = (fmask >> e) & 1u;
    T accb[3];
    if (has_face) {
      if (p < kPre) {
#pragma unroll
        for (int d = 0; d < 3; ++d)
          accb[d] = p == 0 ? pre[0][d] : (p == 1 ? pre[1][d] : pre[2][d]);
      } else {
        const S* a = reinterpret_cast<const S*>(ubar + 12 * (uint64_t)rows[e]);
        accb[0] = (T)a[0];
        accb[1] = (T)a[1];
        accb[2] = (T)a[2];
      }
      ++p;
    }
    T n[3];
    raw_normal(own, e, n);
    if (!(sq3(n[0], n[1], n[2]) > T(0))) continue;  // excluded from the average
    if (!has_face) edge_slot_adjoint(G, act, wprefix, i, j, k, e, ubar_e, accb);
    hermite_vjp(own, e, accb, zero3, Sb);
  }
}

// Phase 2 for one row e of a cell (kind: its kind from phase 1), for one lane of a cell's lane
// group: the same terms as backward_faces adds for that row.
template <typename T, typename S>
__device__ __forceinline__ void backward_face_row(const DcGridDev& G,
                                                  const uint64_t* __restrict__ act,
                                                  const uint32_t* __restrict__ wprefix, int i,
                                                  int j, int k, const Cell<T>& own, int e,
                                                  uint32_t kind, const float* __restrict__ ubar,
                                                  const float* __restrict__ ubar_e, uint64_t nf,
                                                  T Sb[8]) {
  if ((uint64_t)kind >= nf && kind != kRowEdgeSlot) return;  // the own normal: no face term
  T accb[3];
  if (kind < kRowEdgeSlot) {
    const S* a = reinterpret_cast<const S*>(ubar + 12 * (uint64_t)kind);
    accb[0] = (T)a[0];
    accb[1] = (T)a[1];
    accb[2] = (T)a[2];
  }
  T n[3];
  raw_normal(own, e, n);
  if (!(sq3(n[0], n[1], n[2]) > T(0))) return;  // excluded from the average
  if (kind == kRowEdgeSlot) edge_slot_adjoint(G, act, wprefix, i, j, k, e, ubar_e, accb);
  const T zero3[3] = {T(0), T(0), T(0)};
  hermite_vjp(own, e, accb, zero3, Sb);
}

}  // namespace dccell
