// dc_oracle.cpp — CPU restatement of the reference CpuDualContouringBackend (plain dual
// contouring of a dense FP32 SDF grid), used ONLY as the parity checker.
//
// TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load this library. The product (phanty-sdf-cuda_amd/) never links, loads or calls it.
//
// The reference path cannot be compiled here (it needs Eigen + libigl, fetched over the network by
// cmake/libigl.cmake:5-11), so this file restates its algorithm in plain C++17 with the same
// operation order. Build with -ffp-contract=off (x86-64 SSE2, no FMA) so that every FP32 and FP64
// operation rounds exactly where the reference's source rounds.
//
// What follows the reference, step by step (file:line in the reference tree):
//   positions   GV = ox + i*vx in FP32, widened to double      src/cpp/dc_backend.cpp:54-68
//   has_vertex  any corner < 0 and any corner > 0               src/cpp/Cell.cpp:57-77
//   hermite     fa*fb > 0 skip; t = fa/(fa-fb+1e-16); lerp;      src/cpp/Cell.cpp:162-212, 49-52
//               trilinear in-cell gradient, normalized           src/cpp/Cell.cpp:80-155
//   averaging   4-cell normal averaging per interior grid edge   src/cpp/hermite_update.cpp:171-438
//   solve       M = A^T A, r = A^T b, truncated (sigma > 0.01)   src/cpp/Cell.cpp:284-361, 586-600
//               pseudo-inverse about the centroid
//   mesh        vertex ids in (k,j,i) order; quads per axis in   src/cpp/contouring.cpp:486-741
//               (i,j,k) order, k fastest; orientation from
//               normalized central differences                  src/cpp/contouring.cpp:141-207
//   output      float vertices, int quads                        src/cpp/dc_backend.cpp:79-92
//
// Deliberate differences (none changes the mesh):
//   * The per-corner gradientAt() calls of generate_cells (contouring.cpp:243-250) are dead work on
//     the DC path (only the viewer reads cornerNormals) and are skipped.
//   * The second average_hermite_normals() after the solve (contouring.cpp:1242) does not affect
//     V or F and is skipped.
//   * Memory is lean: hermite data is kept only for active cells (the reference keeps ~864 B for
//     every cell, twice), so 512^3 grids fit.
//   * Eigen's JacobiSVD of the 3x3 normal matrix is replaced by a cyclic Jacobi eigen-solver of the
//     symmetric PSD matrix (Golub & Van Loan, "Matrix Computations", sec. 8.5, cyclic Jacobi).
//     For M symmetric PSD, V diag(1/sigma) U^T restricted to sigma > 0.01 equals
//     sum_{lambda_i > 0.01} q_i q_i^T / lambda_i, so the vertex agrees to rounding; only an
//     eigenvalue within rounding of 0.01 could differ. Reduction orders inside Eigen (products,
//     norms) are restated as left-to-right sums except for 3-vectors, where Eigen's SSE2 reduction
//     computes (x*x + y*y) + z*z; differences are at the 1e-16 relative level.
//
// The FP64 analytic adjoint (oracle_dc_backward) differentiates this forward at fixed topology and
// fixed truncation rank; it is the reference gradient for the GPU adjoint.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include <omp.h>

namespace {

// Cell corner order (contouring.cpp:232-241) and edge pairs (include/Cell.h:26-30).
const int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                           {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
const int kEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                          {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
// get_face_normals_from_edge_idx (Cell.cpp:26-45); face ids 0:+X 1:-X 2:+Y 3:-Y 4:+Z 5:-Z.
const int kEdgeFaces[12][2] = {{3, 5}, {0, 5}, {2, 5}, {1, 5}, {3, 4}, {0, 4},
                               {2, 4}, {1, 4}, {1, 3}, {0, 3}, {0, 2}, {1, 2}};
const int kDirs[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

struct Grid {
  int nx, ny, nz;
  float vx, vy, vz, ox, oy, oz;
  const double* val;  // the reference widens values to double (dc_backend.cpp:62)
  int cx() const { return nx - 1; }
  int cy() const { return ny - 1; }
  int cz() const { return nz - 1; }
  int64_t sidx(int i, int j, int k) const { return i + (int64_t)nx * (j + (int64_t)ny * k); }
  int64_t cidx(int i, int j, int k) const { return i + (int64_t)cx() * (j + (int64_t)cy() * k); }
  double S(int i, int j, int k) const { return val[sidx(i, j, k)]; }
};

// dc_backend.cpp:63-65: `grid.ox + i * grid.vx` is an FP32 multiply then an FP32 add.
inline double coord(float o, int i, float v) {
  float prod = (float)i * v;
  float sum = o + prod;
  return (double)sum;
}

inline double sqnorm3(const double v[3]) { return (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]; }
inline double dot3(const double a[3], const double b[3]) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
// Eigen normalize(): divide by sqrt(squaredNorm) when it is > 0.
inline void normalize3(double v[3]) {
  double z = sqnorm3(v);
  if (z > 0.0) {
    double s = std::sqrt(z);
    v[0] = v[0] / s;
    v[1] = v[1] / s;
    v[2] = v[2] / s;
  }
}
inline double clamp01(double t) { return std::max(0.0, std::min(1.0, t)); }

struct CellData {
  double S[8];
  double C[8][3];
};

void load_cell(const Grid& g, int i, int j, int k, CellData& c) {
  for (int q = 0; q < 8; ++q) {
    int gi = i + kCorner[q][0], gj = j + kCorner[q][1], gk = k + kCorner[q][2];
    c.S[q] = g.S(gi, gj, gk);
    c.C[q][0] = coord(g.ox, gi, g.vx);
    c.C[q][1] = coord(g.oy, gj, g.vy);
    c.C[q][2] = coord(g.oz, gk, g.vz);
  }
}

// Cell.cpp:80-155 gradientAtPointUsingTrilinear.
void trilinear_grad(const CellData& c, const double p[3], double g[3]) {
  const double S000 = c.S[0], S100 = c.S[1], S110 = c.S[2], S010 = c.S[3];
  const double S001 = c.S[4], S101 = c.S[5], S111 = c.S[6], S011 = c.S[7];
  const double x0 = c.C[0][0], x1 = c.C[1][0];
  const double y0 = c.C[0][1], y1 = c.C[3][1];
  const double z0 = c.C[0][2], z1 = c.C[4][2];
  const double dx = x1 - x0, dy = y1 - y0, dz = z1 - z0;
  const double eps = 1e-12;
  double u = 0.5, v = 0.5, w = 0.5;
  if (std::abs(dx) > eps) u = (p[0] - x0) / dx;
  if (std::abs(dy) > eps) v = (p[1] - y0) / dy;
  if (std::abs(dz) > eps) w = (p[2] - z0) / dz;
  u = clamp01(u);
  v = clamp01(v);
  w = clamp01(w);
  const double omu = 1.0 - u, omv = 1.0 - v, omw = 1.0 - w;
  double df_du = (S100 - S000) * omv * omw + (S110 - S010) * v * omw + (S101 - S001) * omv * w +
                 (S111 - S011) * v * w;
  double df_dv = (S010 - S000) * omu * omw + (S110 - S100) * u * omw + (S011 - S001) * omu * w +
                 (S111 - S101) * u * w;
  double df_dw = (S001 - S000) * omu * omv + (S101 - S100) * u * omv + (S011 - S010) * omu * v +
                 (S111 - S110) * u * v;
  g[0] = 0.0;
  g[1] = 0.0;
  g[2] = 0.0;
  if (std::abs(dx) > eps) g[0] = df_du / dx;
  if (std::abs(dy) > eps) g[1] = df_dv / dy;
  if (std::abs(dz) > eps) g[2] = df_dw / dz;
}

// Cell.cpp:185-209 for one edge: returns false when fa*fb > 0 (no hermite sample).
bool hermite_edge(const CellData& c, int e, double p[3], double n[3]) {
  const int a = kEdge[e][0], b = kEdge[e][1];
  const double fa = c.S[a], fb = c.S[b];
  if (fa * fb > 0.0) return false;
  const double t = fa / (fa - fb + 1e-16);
  for (int d = 0; d < 3; ++d) p[d] = (1.0 - t) * c.C[a][d] + t * c.C[b][d];
  trilinear_grad(c, p, n);
  normalize3(n);
  return true;
}

// Cyclic Jacobi eigen-decomposition of a symmetric 3x3 matrix (Golub & Van Loan sec. 8.5).
// On return lam[i] are the eigenvalues and Q[:,i] the eigenvectors. The same operation sequence
// is used by the GPU kernel so the two agree bit for bit.
void sym3_eig(const double Min[3][3], double lam[3], double Q[3][3]) {
  double a[3][3];
  for (int r = 0; r < 3; ++r)
    for (int s = 0; s < 3; ++s) {
      a[r][s] = Min[r][s];
      Q[r][s] = (r == s) ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 24; ++sweep) {
    const double off = (a[0][1] * a[0][1] + a[0][2] * a[0][2]) + a[1][2] * a[1][2];
    const double dia = (a[0][0] * a[0][0] + a[1][1] * a[1][1]) + a[2][2] * a[2][2];
    if (!(off > 1e-36 * dia)) break;
    for (int pr = 0; pr < 3; ++pr) {
      const int p = (pr == 2) ? 1 : 0;
      const int q = (pr == 0) ? 1 : 2;
      const int r = 3 - p - q;
      const double apq = a[p][q];
      if (apq == 0.0) continue;
      const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
      const double at = std::abs(theta);
      double t;
      if (at > 1e150)
        t = 0.5 / theta;
      else
        t = 1.0 / (at + std::sqrt(theta * theta + 1.0));
      if (theta < 0.0) t = -t;
      const double c = 1.0 / std::sqrt(t * t + 1.0);
      const double s = t * c;
      const double arp = a[r][p], arq = a[r][q];
      a[p][p] = a[p][p] - t * apq;
      a[q][q] = a[q][q] + t * apq;
      a[p][q] = 0.0;
      a[q][p] = 0.0;
      const double nrp = c * arp - s * arq;
      const double nrq = s * arp + c * arq;
      a[r][p] = nrp;
      a[p][r] = nrp;
      a[r][q] = nrq;
      a[q][r] = nrq;
      for (int row = 0; row < 3; ++row) {
        const double qp = Q[row][p], qq = Q[row][q];
        Q[row][p] = c * qp - s * qq;
        Q[row][q] = s * qp + c * qq;
      }
    }
  }
  lam[0] = a[0][0];
  lam[1] = a[1][1];
  lam[2] = a[2][2];
}

const double kSvdThreshold = 0.01;  // Cell.cpp:314 (hard-coded for the DC path)

struct CellQef {
  int m;
  uint16_t mask;         // edges with hermite samples
  double p[12][3];       // hermite points
  double n[12][3];       // normals (before the solve's re-normalization)
};

// Cell.cpp:284-361 + getCentroid (586-600): returns the vertex.
struct SolveTrace {  // intermediate values kept for the adjoint
  double rows[12][3];
  double b[12];
  int e[12];
  int m;
  double M[3][3], r[3], c[3], g[3], lam[3], Q[3][3];
};

void solve_cell(const CellQef& h, double x[3], SolveTrace* tr) {
  double A[12][3], b[12];
  int ids[12];
  int m = 0;
  for (int e = 0; e < 12; ++e) {
    if (!(h.mask & (1u << e))) continue;
    double n[3] = {h.n[e][0], h.n[e][1], h.n[e][2]};
    if (std::sqrt(sqnorm3(n)) > 0) normalize3(n);
    A[m][0] = n[0];
    A[m][1] = n[1];
    A[m][2] = n[2];
    b[m] = dot3(n, h.p[e]);
    ids[m] = e;
    ++m;
  }
  double M[3][3], r[3];
  for (int u = 0; u < 3; ++u) {
    for (int v = 0; v < 3; ++v) {
      double s = 0.0;
      for (int k = 0; k < m; ++k) s = s + A[k][u] * A[k][v];
      M[u][v] = s;
    }
    double s = 0.0;
    for (int k = 0; k < m; ++k) s = s + A[k][u] * b[k];
    r[u] = s;
  }
  double lam[3], Q[3][3];
  sym3_eig(M, lam, Q);
  double c[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < m; ++k) {
    const int e = ids[k];
    c[0] = c[0] + h.p[e][0];
    c[1] = c[1] + h.p[e][1];
    c[2] = c[2] + h.p[e][2];
  }
  const double dm = (double)m;
  c[0] = c[0] / dm;
  c[1] = c[1] / dm;
  c[2] = c[2] / dm;
  double g[3];
  for (int u = 0; u < 3; ++u) g[u] = r[u] - ((M[u][0] * c[0] + M[u][1] * c[1]) + M[u][2] * c[2]);
  double y[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < 3; ++i) {
    if (!(std::abs(lam[i]) > kSvdThreshold)) continue;
    const double qg = (Q[0][i] * g[0] + Q[1][i] * g[1]) + Q[2][i] * g[2];
    const double coef = qg / lam[i];
    y[0] = y[0] + coef * Q[0][i];
    y[1] = y[1] + coef * Q[1][i];
    y[2] = y[2] + coef * Q[2][i];
  }
  x[0] = c[0] + y[0];
  x[1] = c[1] + y[1];
  x[2] = c[2] + y[2];
  if (tr) {
    tr->m = m;
    for (int k = 0; k < m; ++k) {
      tr->rows[k][0] = A[k][0];
      tr->rows[k][1] = A[k][1];
      tr->rows[k][2] = A[k][2];
      tr->b[k] = b[k];
      tr->e[k] = ids[k];
    }
    for (int u = 0; u < 3; ++u) {
      tr->r[u] = r[u];
      tr->c[u] = c[u];
      tr->g[u] = g[u];
      tr->lam[u] = lam[u];
      for (int v = 0; v < 3; ++v) {
        tr->M[u][v] = M[u][v];
        tr->Q[u][v] = Q[u][v];
      }
    }
  }
}

// contouring.cpp:141-207 gradientAt: central differences with clamped neighbours, normalized.
void grid_gradient(const Grid& g, int i, int j, int k, double out[3]) {
  auto clampi = [](int v, int lo, int hi) { return std::max(lo, std::min(hi, v)); };
  const int im = clampi(i - 1, 0, g.nx - 1), ip = clampi(i + 1, 0, g.nx - 1);
  const int jm = clampi(j - 1, 0, g.ny - 1), jp = clampi(j + 1, 0, g.ny - 1);
  const int km = clampi(k - 1, 0, g.nz - 1), kp = clampi(k + 1, 0, g.nz - 1);
  out[0] = out[1] = out[2] = 0.0;
  const double dx = coord(g.ox, ip, g.vx) - coord(g.ox, im, g.vx);
  if (std::abs(dx) > 1e-12) out[0] = (g.S(ip, j, k) - g.S(im, j, k)) / dx;
  const double dy = coord(g.oy, jp, g.vy) - coord(g.oy, jm, g.vy);
  if (std::abs(dy) > 1e-12) out[1] = (g.S(i, jp, k) - g.S(i, jm, k)) / dy;
  const double dz = coord(g.oz, kp, g.vz) - coord(g.oz, km, g.vz);
  if (std::abs(dz) > 1e-12) out[2] = (g.S(i, j, kp) - g.S(i, j, km)) / dz;
  const double nrm = std::sqrt(sqnorm3(out));
  if (nrm > 0.0) {
    out[0] = out[0] / nrm;
    out[1] = out[1] / nrm;
    out[2] = out[2] / nrm;
  }
}

inline void cross3(const double a[3], const double b[3], double o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// Extraction state, exposed through a small C ABI for the tests (ctypes).
// ---------------------------------------------------------------------------------------------
struct OracleResult {
  Grid g;
  std::vector<double> values;  // private copy of the input, widened (the adjoint re-reads it)
  std::vector<int32_t> vid;   // cell -> vertex id, -1 if inactive
  std::vector<int64_t> active;  // active cell linear ids, vertex order
  std::vector<CellQef> herm;  // per active cell (after normal averaging)
  std::vector<CellQef> herm_raw;  // per active cell, before averaging
  std::vector<double> V;      // A x 3 FP64 vertices
  std::vector<int32_t> F;     // quads
  std::vector<double> dots;   // per quad: quad_n . grid_n (0 when the test was skipped)
  int64_t averaged_edges = 0;
};

namespace {

// hermite_update.cpp:41-91 get_edge_idx_from_grid_vertices.
int edge_idx_in_cell(int ci, int cj, int ck, int i0, int j0, int k0, int i1, int j1, int k1) {
  const int di0 = i0 - ci, dj0 = j0 - cj, dk0 = k0 - ck;
  const int di1 = i1 - ci, dj1 = j1 - cj, dk1 = k1 - ck;
  for (int e = 0; e < 12; ++e) {
    const int* A = kCorner[kEdge[e][0]];
    const int* B = kCorner[kEdge[e][1]];
    const bool fwd = di0 == A[0] && dj0 == A[1] && dk0 == A[2] && di1 == B[0] && dj1 == B[1] &&
                     dk1 == B[2];
    const bool bwd = di0 == B[0] && dj0 == B[1] && dk0 == B[2] && di1 == A[0] && dj1 == A[1] &&
                     dk1 == A[2];
    if (fwd || bwd) return e;
  }
  return -1;
}

// hermite_update.cpp:171-332 process_edge + the averaging loop at :406-437, applied at once
// (each (cell, local edge) slot belongs to exactly one grid edge, so order does not matter).
// Returns 1 when the edge's normals were averaged (for the count), else 0.
int process_edge(OracleResult& R, int i0, int j0, int k0, int i1, int j1, int k1) {
  const Grid& g = R.g;
  auto inside = [&](int a, int b, int c) {
    return a > 0 && a < g.nx - 1 && b > 0 && b < g.ny - 1 && c > 0 && c < g.nz - 1;
  };
  if (!inside(i0, j0, k0) && !inside(i1, j1, k1)) return 0;
  // With one endpoint strictly interior, the cell whose min corner is (i0,j0,k0) exists.
  const int64_t cell = g.cidx(i0, j0, k0);
  const int v0 = R.vid[cell];
  if (v0 < 0) return 0;  // inactive cell: no hermite data
  const int e0 = edge_idx_in_cell(i0, j0, k0, i0, j0, k0, i1, j1, k1);
  if (!(R.herm[v0].mask & (1u << e0))) return 0;
  const int* d1 = kDirs[kEdgeFaces[e0][0]];
  const int* d2 = kDirs[kEdgeFaces[e0][1]];
  const int cc[4][3] = {{i0, j0, k0},
                        {i0 + d1[0], j0 + d1[1], k0 + d1[2]},
                        {i0 + d2[0], j0 + d2[1], k0 + d2[2]},
                        {i0 + d1[0] + d2[0], j0 + d1[1] + d2[1], k0 + d1[2] + d2[2]}};
  int vids[4], eids[4];
  for (int q = 0; q < 4; ++q) {
    vids[q] = R.vid[g.cidx(cc[q][0], cc[q][1], cc[q][2])];
    if (vids[q] < 0) return 0;
  }
  for (int q = 0; q < 4; ++q)
    eids[q] = edge_idx_in_cell(cc[q][0], cc[q][1], cc[q][2], i0, j0, k0, i1, j1, k1);
  double acc[3] = {0.0, 0.0, 0.0};
  for (int q = 0; q < 4; ++q) {
    const double* n = R.herm_raw[vids[q]].n[eids[q]];
    if (std::sqrt(sqnorm3(n)) > 0) {
      acc[0] = acc[0] + n[0];
      acc[1] = acc[1] + n[1];
      acc[2] = acc[2] + n[2];
    }
  }
  normalize3(acc);
  for (int q = 0; q < 4; ++q) {
    double* n = R.herm[vids[q]].n[eids[q]];
    n[0] = acc[0];
    n[1] = acc[1];
    n[2] = acc[2];
  }
  return 1;
}

// The loops below run over OpenMP threads (oracle_set_threads; one thread reproduces the
// reference's single-threaded run). Nothing they compute depends on the thread count: vertex ids
// come from per-layer counts and a prefix over the layers, every (cell, local edge) normal slot is
// written by the one grid edge it belongs to, and per-cell work is independent.
void run_forward(OracleResult& R) {
  const Grid& g = R.g;
  const int cx = g.cx(), cy = g.cy(), cz = g.cz();
  const int64_t ncell = (int64_t)cx * cy * cz;
  const int64_t layer = (int64_t)cx * cy;
  R.vid.assign(ncell, -1);
  // generate_cells: has_vertex per cell (Cell.cpp:66-76), vertex ids in (k,j,i) order
  // (contouring.cpp:509-524): flags and per-layer counts, then ids from the layers' prefix.
  std::vector<int64_t> layer_base(cz + 1, 0);
#pragma omp parallel for schedule(dynamic, 1)
  for (int k = 0; k < cz; ++k) {
    int64_t n = 0;
    for (int j = 0; j < cy; ++j)
      for (int i = 0; i < cx; ++i) {
        bool neg = false, pos = false;
        for (int q = 0; q < 8; ++q) {
          const double s = g.S(i + kCorner[q][0], j + kCorner[q][1], k + kCorner[q][2]);
          if (s < 0) neg = true;
          if (s > 0) pos = true;
        }
        if (neg && pos) {
          R.vid[g.cidx(i, j, k)] = 0;
          ++n;
        }
      }
    layer_base[k + 1] = n;
  }
  for (int k = 0; k < cz; ++k) layer_base[k + 1] += layer_base[k];
  const int64_t A = layer_base[cz];
  R.active.assign(A, 0);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < cz; ++k) {
    int64_t v = layer_base[k];
    for (int64_t ci = (int64_t)k * layer; ci < (int64_t)(k + 1) * layer; ++ci)
      if (R.vid[ci] == 0) {
        R.vid[ci] = (int32_t)v;
        R.active[v++] = ci;
      }
  }
  R.herm.resize(A);
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < A; ++v) {
    const int64_t ci = R.active[v];
    const int i = (int)(ci % cx), j = (int)((ci / cx) % cy), k = (int)(ci / ((int64_t)cx * cy));
    CellData c;
    load_cell(g, i, j, k, c);
    CellQef& h = R.herm[v];
    h.mask = 0;
    h.m = 0;
    for (int e = 0; e < 12; ++e) {
      if (hermite_edge(c, e, h.p[e], h.n[e])) {
        h.mask |= (uint16_t)(1u << e);
        h.m++;
      }
    }
  }
  R.herm_raw = R.herm;
  // average_hermite_normals (hermite_update.cpp:334-438): grid points in (k,j,i) order,
  // X, Y, Z edge per point.
  int64_t averaged = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : averaged)
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        if (i < g.nx - 1) averaged += process_edge(R, i, j, k, i + 1, j, k);
        if (j < g.ny - 1) averaged += process_edge(R, i, j, k, i, j + 1, k);
        if (k < g.nz - 1) averaged += process_edge(R, i, j, k, i, j, k + 1);
      }
  R.averaged_edges = averaged;
  // Cell::update for every active cell.
  R.V.resize(3 * A);
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < A; ++v) solve_cell(R.herm[v], &R.V[3 * v], nullptr);
  // extract_mesh_from_cells (contouring.cpp:607-729).
  auto cellv = [&](int i, int j, int k) -> int {
    if (i < 0 || i >= cx || j < 0 || j >= cy || k < 0 || k >= cz) return -1;
    return R.vid[g.cidx(i, j, k)];
  };
  auto emit = [&](int c0, int c1, int c2, int c3, int gi0, int gj0, int gk0, int gi1, int gj1,
                  int gk1) {
    double g0[3], g1[3];
    grid_gradient(g, gi0, gj0, gk0, g0);
    grid_gradient(g, gi1, gj1, gk1, g1);
    const double gn[3] = {g0[0] + g1[0], g0[1] + g1[1], g0[2] + g1[2]};
    const double* pa = &R.V[3 * (int64_t)c0];
    const double* pb = &R.V[3 * (int64_t)c1];
    const double* pc = &R.V[3 * (int64_t)c2];
    const double* pd = &R.V[3 * (int64_t)c3];
    const double ba[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
    const double da[3] = {pd[0] - pa[0], pd[1] - pa[1], pd[2] - pa[2]};
    double qn[3];
    cross3(ba, da, qn);
    if (sqnorm3(qn) < 1e-24) {
      const double ca[3] = {pc[0] - pa[0], pc[1] - pa[1], pc[2] - pa[2]};
      cross3(ca, da, qn);
    }
    double dt = 0.0;
    if (sqnorm3(gn) > 0 && sqnorm3(qn) > 0) {
      dt = dot3(qn, gn);
      if (dt < 0.0) std::swap(c1, c3);
    }
    R.F.push_back(c0);
    R.F.push_back(c1);
    R.F.push_back(c2);
    R.F.push_back(c3);
    R.dots.push_back(dt);
  };
  auto sign_change = [&](int a0, int b0, int c0, int a1, int b1, int c1) {
    return g.S(a0, b0, c0) * g.S(a1, b1, c1) < 0.0;
  };
  // The reference visits each axis's edges in (i, j, k) order, k innermost (contouring.cpp:
  // 607-729), a stride of nx*ny samples per step. The sign-changing edges are found here in
  // memory order (k, j, i) and sorted into the reference's (i, j, k) order before emission, which
  // is the same sequence of quads at a fraction of the cache misses on large grids.
  auto edges_in_order = [&](int axis, int ilo, int jlo, int klo) {
    const int di = axis == 0, dj = axis == 1, dk = axis == 2;
    std::vector<std::vector<uint64_t>> per_k(cz);
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = klo; k < cz; ++k)
      for (int j = jlo; j < cy; ++j)
        for (int i = ilo; i < cx; ++i)
          if (sign_change(i, j, k, i + di, j + dj, k + dk))
            per_k[k].push_back(((uint64_t)i * (uint64_t)cy + (uint64_t)j) * (uint64_t)cz +
                               (uint64_t)k);
    std::vector<uint64_t> keys;
    for (auto& v : per_k) keys.insert(keys.end(), v.begin(), v.end());
    std::sort(keys.begin(), keys.end());
    return keys;
  };
  auto unkey = [&](uint64_t key, int& i, int& j, int& k) {
    k = (int)(key % (uint64_t)cz);
    j = (int)((key / (uint64_t)cz) % (uint64_t)cy);
    i = (int)(key / ((uint64_t)cz * (uint64_t)cy));
  };
  for (uint64_t key : edges_in_order(0, 0, 1, 1)) {  // x: i in [0,cx), j in [1,cy), k in [1,cz)
    int i, j, k;
    unkey(key, i, j, k);
    const int c0 = cellv(i, j - 1, k - 1), c1 = cellv(i, j, k - 1), c2 = cellv(i, j, k),
              c3 = cellv(i, j - 1, k);
    if (c0 < 0 || c1 < 0 || c2 < 0 || c3 < 0) continue;
    emit(c0, c1, c2, c3, i, j, k, i + 1, j, k);
  }
  for (uint64_t key : edges_in_order(1, 1, 0, 1)) {  // y: i in [1,cx), j in [0,cy), k in [1,cz)
    int i, j, k;
    unkey(key, i, j, k);
    const int c0 = cellv(i - 1, j, k - 1), c1 = cellv(i, j, k - 1), c2 = cellv(i, j, k),
              c3 = cellv(i - 1, j, k);
    if (c0 < 0 || c1 < 0 || c2 < 0 || c3 < 0) continue;
    emit(c0, c1, c2, c3, i, j, k, i, j + 1, k);
  }
  for (uint64_t key : edges_in_order(2, 1, 1, 0)) {  // z: i in [1,cx), j in [1,cy), k in [0,cz)
    int i, j, k;
    unkey(key, i, j, k);
    const int c0 = cellv(i - 1, j - 1, k), c1 = cellv(i, j - 1, k), c2 = cellv(i, j, k),
              c3 = cellv(i - 1, j, k);
    if (c0 < 0 || c1 < 0 || c2 < 0 || c3 < 0) continue;
    emit(c0, c1, c2, c3, i, j, k, i, j, k + 1);
  }
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// FP64 analytic adjoint at fixed topology and fixed truncation rank.
//
// Forward chain per active cell v (all in double):
//   t = fa/(fa-fb+1e-16), p = (1-t)pa + t pb, G = trilinear gradient of the cell at p,
//   n_raw = G/|G|, n_fin = normalize(sum of the 4 cells' n_raw) on averaged edges else n_raw,
//   row n = n_fin/|n_fin|, b = n.p, M = sum n n^T, r = sum b n, c = sum p / m, g = r - M c,
//   x = c + P g with P = sum_{lambda_i > 0.01} q_i q_i^T / lambda_i.
// The derivative of P uses the Daleckii-Krein formula for spectral functions: for symmetric E,
// dP = Q (F o (Q^T E Q)) Q^T with F_ij = (f(l_i) - f(l_j)) / (l_i - l_j), F_ii = f'(l_i),
// f(l) = 1/l above the threshold and 0 below (both sides held fixed).
// ---------------------------------------------------------------------------------------------
namespace {

// VJP of the trilinear gradient G(S, p) (Cell.cpp:80-155) with respect to S[8] and p.
void trilinear_grad_vjp(const CellData& c, const double p[3], const double Gb[3], double Sb[8],
                        double pb[3]) {
  const double S000 = c.S[0], S100 = c.S[1], S110 = c.S[2], S010 = c.S[3];
  const double S001 = c.S[4], S101 = c.S[5], S111 = c.S[6], S011 = c.S[7];
  const double x0 = c.C[0][0], x1 = c.C[1][0];
  const double y0 = c.C[0][1], y1 = c.C[3][1];
  const double z0 = c.C[0][2], z1 = c.C[4][2];
  const double dx = x1 - x0, dy = y1 - y0, dz = z1 - z0;
  const double eps = 1e-12;
  const bool hx = std::abs(dx) > eps, hy = std::abs(dy) > eps, hz = std::abs(dz) > eps;
  double u = 0.5, v = 0.5, w = 0.5;
  if (hx) u = (p[0] - x0) / dx;
  if (hy) v = (p[1] - y0) / dy;
  if (hz) w = (p[2] - z0) / dz;
  // clamp01 passes the derivative through inside [0, 1]
  const bool iu = hx && u >= 0.0 && u <= 1.0, iv = hy && v >= 0.0 && v <= 1.0,
             iw = hz && w >= 0.0 && w <= 1.0;
  u = clamp01(u);
  v = clamp01(v);
  w = clamp01(w);
  const double omu = 1.0 - u, omv = 1.0 - v, omw = 1.0 - w;
  // d(df_du) etc. scaled by the chain factor 1/d
  const double a = hx ? Gb[0] / dx : 0.0;  // adjoint of df_du
  const double bb = hy ? Gb[1] / dy : 0.0;  // adjoint of df_dv
  const double cc = hz ? Gb[2] / dz : 0.0;  // adjoint of df_dw
  // df_du terms
  Sb[1] += a * omv * omw; Sb[0] -= a * omv * omw;
  Sb[2] += a * v * omw;   Sb[3] -= a * v * omw;
  Sb[5] += a * omv * w;   Sb[4] -= a * omv * w;
  Sb[6] += a * v * w;     Sb[7] -= a * v * w;
  // df_dv terms
  Sb[3] += bb * omu * omw; Sb[0] -= bb * omu * omw;
  Sb[2] += bb * u * omw;   Sb[1] -= bb * u * omw;
  Sb[7] += bb * omu * w;   Sb[4] -= bb * omu * w;
  Sb[6] += bb * u * w;     Sb[5] -= bb * u * w;
  // df_dw terms
  Sb[4] += cc * omu * omv; Sb[0] -= cc * omu * omv;
  Sb[5] += cc * u * omv;   Sb[1] -= cc * u * omv;
  Sb[7] += cc * omu * v;   Sb[3] -= cc * omu * v;
  Sb[6] += cc * u * v;     Sb[2] -= cc * u * v;
  // derivatives with respect to u, v, w
  const double dfdu_dv = (S110 - S010) - (S100 - S000);
  const double dfdu_dw_ = 0.0;
  (void)dfdu_dw_;
  // d(df_du)/dv, d(df_du)/dw
  const double ddu_dv = ((S110 - S010) - (S100 - S000)) * omw + ((S111 - S011) - (S101 - S001)) * w;
  const double ddu_dw = ((S101 - S001) - (S100 - S000)) * omv + ((S111 - S011) - (S110 - S010)) * v;
  const double ddv_du = ((S110 - S100) - (S010 - S000)) * omw + ((S111 - S101) - (S011 - S001)) * w;
  const double ddv_dw = ((S011 - S001) - (S010 - S000)) * omu + ((S111 - S101) - (S110 - S100)) * u;
  const double ddw_du = ((S101 - S100) - (S001 - S000)) * omv + ((S111 - S110) - (S011 - S010)) * v;
  const double ddw_dv = ((S011 - S010) - (S001 - S000)) * omu + ((S111 - S110) - (S101 - S100)) * u;
  (void)dfdu_dv;
  const double ub = bb * ddv_du + cc * ddw_du;
  const double vb = a * ddu_dv + cc * ddw_dv;
  const double wb = a * ddu_dw + bb * ddv_dw;
  pb[0] += iu ? ub / dx : 0.0;
  pb[1] += iv ? vb / dy : 0.0;
  pb[2] += iw ? wb / dz : 0.0;
}

// VJP of n = normalize(G) (skip when |G| = 0, as Eigen's normalize does).
void normalize_vjp(const double G[3], const double nb[3], double Gb[3]) {
  const double z = sqnorm3(G);
  if (!(z > 0.0)) {
    Gb[0] = nb[0];
    Gb[1] = nb[1];
    Gb[2] = nb[2];
    return;
  }
  const double s = std::sqrt(z);
  const double n[3] = {G[0] / s, G[1] / s, G[2] / s};
  const double d = dot3(n, nb);
  for (int q = 0; q < 3; ++q) Gb[q] = (nb[q] - n[q] * d) / s;
}

// Hermite sample (cell, edge) backward: n_raw-bar and p-bar to the cell's 8 corners.
void hermite_vjp(const CellData& c, int e, const double nb[3], const double pbar_in[3],
                 double Sb[8]) {
  const int a = kEdge[e][0], b = kEdge[e][1];
  const double fa = c.S[a], fb = c.S[b];
  const double D = fa - fb + 1e-16;
  const double t = fa / D;
  double p[3];
  for (int d = 0; d < 3; ++d) p[d] = (1.0 - t) * c.C[a][d] + t * c.C[b][d];
  double G[3];
  trilinear_grad(c, p, G);
  double Gb[3];
  normalize_vjp(G, nb, Gb);
  double pb[3] = {pbar_in[0], pbar_in[1], pbar_in[2]};
  trilinear_grad_vjp(c, p, Gb, Sb, pb);
  const double tb = (pb[0] * (c.C[b][0] - c.C[a][0]) + pb[1] * (c.C[b][1] - c.C[a][1])) +
                    pb[2] * (c.C[b][2] - c.C[a][2]);
  Sb[a] += tb * (1e-16 - fb) / (D * D);
  Sb[b] += tb * fa / (D * D);
}

// The 4 cells around the grid edge carried by local edge e of cell (i,j,k), in the reference's
// process_edge order, with their local edge ids. Returns false if the edge is not averaged.
bool averaged_edge_cells(const OracleResult& R, int i, int j, int k, int e, int cells[4][3],
                         int eids[4]) {
  const Grid& g = R.g;
  const int* A = kCorner[kEdge[e][0]];
  const int* B = kCorner[kEdge[e][1]];
  int P0[3] = {i + A[0], j + A[1], k + A[2]}, P1[3] = {i + B[0], j + B[1], k + B[2]};
  // lower endpoint first (the reference visits each edge from its lower grid point)
  if (P1[0] + P1[1] + P1[2] < P0[0] + P0[1] + P0[2]) std::swap(P0, P1);
  auto inside = [&](const int* P) {
    return P[0] > 0 && P[0] < g.nx - 1 && P[1] > 0 && P[1] < g.ny - 1 && P[2] > 0 &&
           P[2] < g.nz - 1;
  };
  if (!inside(P0) && !inside(P1)) return false;
  const int e0 = edge_idx_in_cell(P0[0], P0[1], P0[2], P0[0], P0[1], P0[2], P1[0], P1[1], P1[2]);
  const int* d1 = kDirs[kEdgeFaces[e0][0]];
  const int* d2 = kDirs[kEdgeFaces[e0][1]];
  const int cc[4][3] = {{P0[0], P0[1], P0[2]},
                        {P0[0] + d1[0], P0[1] + d1[1], P0[2] + d1[2]},
                        {P0[0] + d2[0], P0[1] + d2[1], P0[2] + d2[2]},
                        {P0[0] + d1[0] + d2[0], P0[1] + d1[1] + d2[1], P0[2] + d1[2] + d2[2]}};
  for (int q = 0; q < 4; ++q) {
    if (R.vid[g.cidx(cc[q][0], cc[q][1], cc[q][2])] < 0) return false;
    for (int d = 0; d < 3; ++d) cells[q][d] = cc[q][d];
    eids[q] = edge_idx_in_cell(cc[q][0], cc[q][1], cc[q][2], P0[0], P0[1], P0[2], P1[0], P1[1],
                               P1[2]);
  }
  return true;
}

void add_cell_grad(const Grid& g, int i, int j, int k, const double Sb[8], double* gphi) {
  for (int q = 0; q < 8; ++q)
    gphi[g.sidx(i + kCorner[q][0], j + kCorner[q][1], k + kCorner[q][2])] += Sb[q];
}

void backward_cell(const OracleResult& R, int64_t v, const double xb[3], double* gphi) {
  const Grid& g = R.g;
  const int cx = g.cx(), cy = g.cy();
  const int64_t ci = R.active[v];
  const int i = (int)(ci % cx), j = (int)((ci / cx) % cy), k = (int)(ci / ((int64_t)cx * cy));
  SolveTrace tr;
  double x[3];
  solve_cell(R.herm[v], x, &tr);
  const int m = tr.m;
  // x = c + P g
  double cb[3] = {xb[0], xb[1], xb[2]};
  bool kept[3];
  for (int q = 0; q < 3; ++q) kept[q] = std::abs(tr.lam[q]) > kSvdThreshold;
  double gb[3] = {0, 0, 0};
  for (int q = 0; q < 3; ++q) {
    if (!kept[q]) continue;
    const double qx = (tr.Q[0][q] * xb[0] + tr.Q[1][q] * xb[1]) + tr.Q[2][q] * xb[2];
    for (int d = 0; d < 3; ++d) gb[d] += tr.Q[d][q] * qx / tr.lam[q];
  }
  // P-bar = xb g^T, symmetrized, through Daleckii-Krein
  double Pb[3][3];
  for (int a2 = 0; a2 < 3; ++a2)
    for (int b2 = 0; b2 < 3; ++b2) Pb[a2][b2] = 0.5 * (xb[a2] * tr.g[b2] + tr.g[a2] * xb[b2]);
  double Bm[3][3];
  for (int a2 = 0; a2 < 3; ++a2)
    for (int b2 = 0; b2 < 3; ++b2) {
      double s = 0.0;
      for (int u = 0; u < 3; ++u)
        for (int w = 0; w < 3; ++w) s += tr.Q[u][a2] * Pb[u][w] * tr.Q[w][b2];
      double F;
      const double la = tr.lam[a2], lb = tr.lam[b2];
      if (kept[a2] && kept[b2])
        F = -1.0 / (la * lb);
      else if (!kept[a2] && !kept[b2])
        F = 0.0;
      else
        F = ((kept[a2] ? 1.0 / la : 0.0) - (kept[b2] ? 1.0 / lb : 0.0)) / (la - lb);
      Bm[a2][b2] = F * s;
    }
  double Mb[3][3];
  for (int a2 = 0; a2 < 3; ++a2)
    for (int b2 = 0; b2 < 3; ++b2) {
      double s = 0.0;
      for (int u = 0; u < 3; ++u)
        for (int w = 0; w < 3; ++w) s += tr.Q[a2][u] * Bm[u][w] * tr.Q[b2][w];
      Mb[a2][b2] = s - gb[a2] * tr.c[b2];
    }
  const double rb[3] = {gb[0], gb[1], gb[2]};
  for (int a2 = 0; a2 < 3; ++a2)
    cb[a2] -= (tr.M[a2][0] * gb[0] + tr.M[a2][1] * gb[1]) + tr.M[a2][2] * gb[2];
  CellData own;
  load_cell(g, i, j, k, own);
  double Sb_own[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int kk = 0; kk < m; ++kk) {
    const int e = tr.e[kk];
    const double* n = tr.rows[kk];
    const double* p = R.herm[v].p[e];
    const double bk = tr.b[kk];
    const double bbk = dot3(n, rb);
    double nb[3], pbk[3];
    for (int d = 0; d < 3; ++d) {
      const double Mn = ((Mb[d][0] + Mb[0][d]) * n[0] + (Mb[d][1] + Mb[1][d]) * n[1]) +
                        (Mb[d][2] + Mb[2][d]) * n[2];
      nb[d] = Mn + bk * rb[d] + bbk * p[d];
      pbk[d] = bbk * n[d] + cb[d] / (double)m;
    }
    // row renormalization n = u/|u|, u = n_fin
    const double* u = R.herm[v].n[e];
    double ub[3];
    normalize_vjp(u, nb, ub);
    int cells[4][3], eids[4];
    const double zero3[3] = {0, 0, 0};
    if (averaged_edge_cells(R, i, j, k, e, cells, eids)) {
      double acc[3] = {0, 0, 0};
      bool contrib[4];
      CellData cd[4];
      for (int q = 0; q < 4; ++q) {
        const int vq = R.vid[g.cidx(cells[q][0], cells[q][1], cells[q][2])];
        const double* nq = R.herm_raw[vq].n[eids[q]];
        contrib[q] = std::sqrt(sqnorm3(nq)) > 0;
        if (contrib[q]) {
          acc[0] += nq[0];
          acc[1] += nq[1];
          acc[2] += nq[2];
        }
        load_cell(g, cells[q][0], cells[q][1], cells[q][2], cd[q]);
      }
      double accb[3];
      normalize_vjp(acc, ub, accb);
      for (int q = 0; q < 4; ++q) {
        if (!contrib[q]) continue;
        double Sb[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        hermite_vjp(cd[q], eids[q], accb, zero3, Sb);
        add_cell_grad(g, cells[q][0], cells[q][1], cells[q][2], Sb, gphi);
      }
      // the row's own hermite point (p enters b and c)
      const int a = kEdge[e][0], b = kEdge[e][1];
      const double fa = own.S[a], fb = own.S[b];
      const double D = fa - fb + 1e-16;
      const double tb = (pbk[0] * (own.C[b][0] - own.C[a][0]) + pbk[1] * (own.C[b][1] - own.C[a][1])) +
                        pbk[2] * (own.C[b][2] - own.C[a][2]);
      Sb_own[a] += tb * (1e-16 - fb) / (D * D);
      Sb_own[b] += tb * fa / (D * D);
    } else {
      hermite_vjp(own, e, ub, pbk, Sb_own);
    }
  }
  add_cell_grad(g, i, j, k, Sb_own, gphi);
}

}  // namespace

extern "C" {

OracleResult* oracle_dc_run_f64(int nx, int ny, int nz, float vx, float vy, float vz, float ox,
                                float oy, float oz, const double* values) {
  if (nx < 2 || ny < 2 || nz < 2) return nullptr;
  OracleResult* R = new OracleResult();
  const int64_t ns = (int64_t)nx * ny * nz;
  R->values.assign(values, values + ns);
  R->g = Grid{nx, ny, nz, vx, vy, vz, ox, oy, oz, R->values.data()};
  run_forward(*R);
  return R;
}

// Runs the restated CPU backend on FP32 values (widened to double as dc_backend.cpp:62 does).
// Returns NULL on invalid dimensions.
OracleResult* oracle_dc_run(int nx, int ny, int nz, float vx, float vy, float vz, float ox, float oy,
                            float oz, const float* values) {
  if (nx < 2 || ny < 2 || nz < 2) return nullptr;
  OracleResult* R = new OracleResult();
  const int64_t ns = (int64_t)nx * ny * nz;
  R->values.resize(ns);
  for (int64_t q = 0; q < ns; ++q) R->values[q] = (double)values[q];
  R->g = Grid{nx, ny, nz, vx, vy, vz, ox, oy, oz, R->values.data()};
  run_forward(*R);
  return R;
}

// dL/dphi (dense, N_s doubles, overwritten) for dL/dV given as A x 3 doubles.
void oracle_dc_backward(const OracleResult* R, const double* gV, double* gphi) {
  const int64_t ns = (int64_t)R->g.nx * R->g.ny * R->g.nz;
  std::memset(gphi, 0, ns * sizeof(double));
  for (int64_t v = 0; v < (int64_t)R->active.size(); ++v)
    backward_cell(*R, v, gV + 3 * v, gphi);
}

void oracle_dc_free(OracleResult* R) { delete R; }

// Threads of the forward's loops (1: the reference's single-threaded run).
void oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

// triangulate_v1 (contouring.cpp:88-113) with orient_triangles_for_quad (contouring.cpp:40-86):
// quad f = (a, b, c, d) -> triangles 2f = (a, b, c) and 2f+1 = (a, c, d). The quad normal is
// (pb - pa) x (pd - pa), or (pc - pa) x (pd - pa) when the first has squared length < 1e-24; a
// quad whose normal is still that short keeps both triangles as they are. Otherwise a triangle
// whose normal (p2 - p1) x (p3 - p1) has a negative dot with the quad normal has its second and
// third vertices swapped. V: nv x 3 doubles; T: 2 nf x 3 int32.
void oracle_triangulate(const int32_t* F, int64_t nf, const double* V, int64_t nv, int32_t* T) {
  (void)nv;
  auto P = [&](int32_t q, double out[3]) {
    out[0] = V[3 * (int64_t)q];
    out[1] = V[3 * (int64_t)q + 1];
    out[2] = V[3 * (int64_t)q + 2];
  };
  auto cross = [](const double u[3], const double w[3], double o[3]) {
    o[0] = u[1] * w[2] - u[2] * w[1];
    o[1] = u[2] * w[0] - u[0] * w[2];
    o[2] = u[0] * w[1] - u[1] * w[0];
  };
  auto sub = [](const double u[3], const double w[3], double o[3]) {
    o[0] = u[0] - w[0];
    o[1] = u[1] - w[1];
    o[2] = u[2] - w[2];
  };
  auto dot = [](const double u[3], const double w[3]) {
    return (u[0] * w[0] + u[1] * w[1]) + u[2] * w[2];
  };
  for (int64_t f = 0; f < nf; ++f) {
    const int32_t a = F[4 * f], b = F[4 * f + 1], c = F[4 * f + 2], d = F[4 * f + 3];
    int32_t* t0 = T + 6 * f;
    int32_t* t1 = t0 + 3;
    t0[0] = a; t0[1] = b; t0[2] = c;
    t1[0] = a; t1[1] = c; t1[2] = d;
    double pa[3], pb[3], pc[3], pd[3], u[3], w[3], qn[3];
    P(a, pa); P(b, pb); P(c, pc); P(d, pd);
    sub(pb, pa, u);
    sub(pd, pa, w);
    cross(u, w, qn);
    if (dot(qn, qn) < 1e-24) {
      sub(pc, pa, u);
      cross(u, w, qn);
    }
    if (dot(qn, qn) < 1e-24) continue;
    for (int32_t* t : {t0, t1}) {
      double p1[3], p2[3], p3[3], tn[3];
      P(t[0], p1); P(t[1], p2); P(t[2], p3);
      sub(p2, p1, u);
      sub(p3, p1, w);
      cross(u, w, tn);
      if (dot(tn, qn) < 0.0) std::swap(t[1], t[2]);
    }
  }
}

void oracle_dc_counts(const OracleResult* R, int64_t* n_vertices, int64_t* n_faces,
                      int64_t* averaged_edges) {
  *n_vertices = (int64_t)R->active.size();
  *n_faces = (int64_t)(R->F.size() / 4);
  if (averaged_edges) *averaged_edges = R->averaged_edges;
}

// Any output pointer may be NULL.
void oracle_dc_copy(const OracleResult* R, float* v32, double* v64, int32_t* faces,
                    int64_t* active_cells, double* face_dots) {
  const size_t A = R->active.size();
  if (v32)
    for (size_t q = 0; q < 3 * A; ++q) v32[q] = (float)R->V[q];  // dc_backend.cpp:80-85
  if (v64) std::memcpy(v64, R->V.data(), 3 * A * sizeof(double));
  if (faces) std::memcpy(faces, R->F.data(), R->F.size() * sizeof(int32_t));
  if (active_cells) std::memcpy(active_cells, R->active.data(), A * sizeof(int64_t));
  if (face_dots) std::memcpy(face_dots, R->dots.data(), R->dots.size() * sizeof(double));
}

// Per active cell: edge mask, hermite points and normals (after averaging when averaged != 0).
void oracle_dc_hermite(const OracleResult* R, int averaged, uint16_t* mask, double* points,
                       double* normals) {
  const std::vector<CellQef>& H = averaged ? R->herm : R->herm_raw;
  for (size_t v = 0; v < H.size(); ++v) {
    mask[v] = H[v].mask;
    std::memcpy(points + 36 * v, H[v].p, 36 * sizeof(double));
    std::memcpy(normals + 36 * v, H[v].n, 36 * sizeof(double));
  }
}

// Exposed pieces for closed-form checks.
void oracle_sym3_eig(const double* M, double* lam, double* Q) {
  double Mm[3][3], Qm[3][3];
  for (int r = 0; r < 3; ++r)
    for (int s = 0; s < 3; ++s) Mm[r][s] = M[3 * r + s];
  sym3_eig(Mm, lam, Qm);
  for (int r = 0; r < 3; ++r)
    for (int s = 0; s < 3; ++s) Q[3 * r + s] = Qm[r][s];
}

// QEF solve of m rows (normals n, points p) exactly as Cell::update does after averaging.
void oracle_qef_solve(int m, const double* normals, const double* points, double* x) {
  CellQef h;
  h.m = m;
  h.mask = 0;
  for (int k = 0; k < m && k < 12; ++k) {
    h.mask |= (uint16_t)(1u << k);
    for (int d = 0; d < 3; ++d) {
      h.n[k][d] = normals[3 * k + d];
      h.p[k][d] = points[3 * k + d];
    }
  }
  solve_cell(h, x, nullptr);
}

}  // extern "C"
