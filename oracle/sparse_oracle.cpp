// sparse_oracle.cpp — TEST INFRASTRUCTURE (the checker, never the product): a CPU restatement of
// the reference's sparse brick dual contouring and its multi-vertex form,
// CudaSparseDualContouringBackend / CudaSparseMvdcDualContouringBackend
// (src/cuda/dc_cuda_sparse.cu:1170-1485, declared include/dc_cuda.h:37-62).
//
// Every floating-point step is FP32 in the reference's operation order (built with
// -ffp-contract=off, so no step is fused), which makes the result a bit-for-bit statement of the
// reference's arithmetic. Order-dependent choices follow the reference: active cells ascend by
// linear cell index (thrust::copy_if over a counting iterator, :1277-1285), vertices follow the
// active cells (two for a split cell, :1316-1327), faces follow the active cells, each cell's x,
// then y, then z edge (:938-1102, exclusive scan :1374). The reference's counters are atomics
// whose totals do not depend on order, so they are plain sums here.
//
// The one intended departure: the reference takes any brick size and any normal mode without
// checks; the entry point returns null for brick_size < 1, a normal mode outside 0..2 or a grid
// under 2 samples per axis (the product rejects the same inputs with an error code).
#include <math.h>
#include <stdint.h>

#include <vector>

namespace {

struct V3 {
  float x, y, z;
};
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 mul(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

struct SGrid {
  const float* f;
  int nx, ny, nz;
  float vx, vy, vz, ox, oy, oz;
  float at(int i, int j, int k) const { return f[(int64_t)i + (int64_t)nx * (j + (int64_t)ny * k)]; }
  int64_t cell(int i, int j, int k) const {
    return (int64_t)i + (int64_t)(nx - 1) * (j + (int64_t)(ny - 1) * k);
  }
};

// corner c of a cell sits at (i + kCorner[c][0], j + kCorner[c][1], k + kCorner[c][2]);
// edge e joins corners kEdge[e][0] and kEdge[e][1] (dc_cuda_sparse.cu:477-486)
const int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                           {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
const int kEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                          {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

// dc_cuda_sparse.cu:151-176: central differences of the grid at a sample, one-sided at the
// border (the neighbour index clamps), divided by the index span times the spacing.
V3 sample_gradient(const SGrid& g, int i, int j, int k) {
  const int im = i > 0 ? i - 1 : 0, ip = i < g.nx - 1 ? i + 1 : g.nx - 1;
  const int jm = j > 0 ? j - 1 : 0, jp = j < g.ny - 1 ? j + 1 : g.ny - 1;
  const int km = k > 0 ? k - 1 : 0, kp = k < g.nz - 1 ? k + 1 : g.nz - 1;
  const float fim = g.at(im, j, k), fip = g.at(ip, j, k);
  const float fjm = g.at(i, jm, k), fjp = g.at(i, jp, k);
  const float fkm = g.at(i, j, km), fkp = g.at(i, j, kp);
  V3 r = v3(0.f, 0.f, 0.f);
  const float dx = (float)(ip - im) * g.vx, dy = (float)(jp - jm) * g.vy,
              dz = (float)(kp - km) * g.vz;
  if (dx > 1e-12f) r.x = (fip - fim) / dx;
  if (dy > 1e-12f) r.y = (fjp - fjm) / dy;
  if (dz > 1e-12f) r.z = (fkp - fkm) / dz;
  return r;
}

// dc_cuda_sparse.cu:90-149: gradient of the cell's trilinear interpolant at p (p's cell
// coordinates clamped to [0, 1]).
V3 trilinear_gradient(const float f[8], const V3 c[8], V3 p) {
  const float dx = c[1].x - c[0].x, dy = c[3].y - c[0].y, dz = c[4].z - c[0].z;
  float u = 0.5f, v = 0.5f, w = 0.5f;
  if (fabsf(dx) > 1e-12f) u = (p.x - c[0].x) / dx;
  if (fabsf(dy) > 1e-12f) v = (p.y - c[0].y) / dy;
  if (fabsf(dz) > 1e-12f) w = (p.z - c[0].z) / dz;
  u = fmaxf(0.0f, fminf(1.0f, u));
  v = fmaxf(0.0f, fminf(1.0f, v));
  w = fmaxf(0.0f, fminf(1.0f, w));
  const float ou = 1.0f - u, ov = 1.0f - v, ow = 1.0f - w;
  // f index: 0 (000) 1 (100) 2 (110) 3 (010) 4 (001) 5 (101) 6 (111) 7 (011)
  const float gu = (f[1] - f[0]) * ov * ow + (f[2] - f[3]) * v * ow + (f[5] - f[4]) * ov * w +
                   (f[6] - f[7]) * v * w;
  const float gv = (f[3] - f[0]) * ou * ow + (f[2] - f[1]) * u * ow + (f[7] - f[4]) * ou * w +
                   (f[6] - f[5]) * u * w;
  const float gw = (f[4] - f[0]) * ou * ov + (f[5] - f[1]) * u * ov + (f[7] - f[3]) * ou * v +
                   (f[6] - f[2]) * u * v;
  V3 r = v3(0.f, 0.f, 0.f);
  if (fabsf(dx) > 1e-12f) r.x = gu / dx;
  if (fabsf(dy) > 1e-12f) r.y = gv / dy;
  if (fabsf(dz) > 1e-12f) r.z = gw / dz;
  return r;
}

// dc_cuda_sparse.cu:310-319: a same-sign edge whose two samples both lie within 0.45 vx of the
// surface and whose end gradients point against each other (a sheet thinner than the cell).
bool thin_edge(const SGrid& g, float fa, float fb, const int a[3], const int b[3]) {
  if (fa * fb <= 0.0f) return false;
  const float lim = 0.45f * g.vx;
  if (fabsf(fa) > lim || fabsf(fb) > lim) return false;
  const V3 ga = sample_gradient(g, a[0], a[1], a[2]);
  const V3 gb = sample_gradient(g, b[0], b[1], b[2]);
  return dot(ga, gb) < -0.3f;
}

// dc_cuda_sparse.cu:178-269: cyclic Jacobi on the symmetric 3x3 B (largest upper off-diagonal
// entry each sweep, at most 15 rotations, stop below 1e-7), then y = V S+ V^T d with singular
// values at or under the threshold dropped. *singular: every singular value dropped.
void jacobi_pinv(const float B[9], const float d[3], float thr, float y[3], bool* singular) {
  float V[9] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
  float A[9];
  for (int i = 0; i < 9; ++i) A[i] = B[i];
  for (int it = 0; it < 15; ++it) {
    int p = 0, q = 1;
    float big = fabsf(A[1]);
    if (fabsf(A[2]) > big) {
      p = 0;
      q = 2;
      big = fabsf(A[2]);
    }
    if (fabsf(A[5]) > big) {
      p = 1;
      q = 2;
      big = fabsf(A[5]);
    }
    if (big < 1e-7f) break;
    const float app = A[p * 3 + p], aqq = A[q * 3 + q], apq = A[p * 3 + q];
    const float phi = 0.5f * (aqq - app) / apq;
    float t = 1.0f / (fabsf(phi) + sqrtf(phi * phi + 1.0f));
    if (phi < 0.0f) t = -t;
    const float c = 1.0f / sqrtf(t * t + 1.0f);
    const float s = t * c;
    for (int r = 0; r < 3; ++r) {
      if (r == p || r == q) continue;
      const float arp = A[r * 3 + p], arq = A[r * 3 + q];
      A[r * 3 + p] = A[p * 3 + r] = c * arp - s * arq;
      A[r * 3 + q] = A[q * 3 + r] = s * arp + c * arq;
    }
    A[p * 3 + p] = app - t * apq;
    A[q * 3 + q] = aqq + t * apq;
    A[p * 3 + q] = A[q * 3 + p] = 0.0f;
    for (int r = 0; r < 3; ++r) {
      const float vrp = V[r * 3 + p], vrq = V[r * 3 + q];
      V[r * 3 + p] = c * vrp - s * vrq;
      V[r * 3 + q] = s * vrp + c * vrq;
    }
  }
  const float S[3] = {A[0], A[4], A[8]};
  float Si[3];
  int kept = 0;
  for (int i = 0; i < 3; ++i) {
    if (S[i] > thr) {
      Si[i] = 1.0f / S[i];
      ++kept;
    } else {
      Si[i] = 0.0f;
    }
  }
  *singular = kept == 0;
  float w[3];
  w[0] = V[0] * d[0] + V[3] * d[1] + V[6] * d[2];
  w[1] = V[1] * d[0] + V[4] * d[1] + V[7] * d[2];
  w[2] = V[2] * d[0] + V[5] * d[1] + V[8] * d[2];
  w[0] *= Si[0];
  w[1] *= Si[1];
  w[2] *= Si[2];
  y[0] = V[0] * w[0] + V[1] * w[1] + V[2] * w[2];
  y[1] = V[3] * w[0] + V[4] * w[1] + V[5] * w[2];
  y[2] = V[6] * w[0] + V[7] * w[1] + V[8] * w[2];
}

struct Counters {
  int64_t fallbacks = 0, clamps = 0, invalid = 0, ambiguous = 0, multi = 0, one = 0, two = 0,
          split_rejections = 0, bad_qef = 0, faces_skipped = 0;
};

bool finite3(V3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }

// dc_cuda_sparse.cu:380-443: the QEF of the entries selected by `take` (bit i: entry i), solved
// about their mass point; the mass point when the solve drops every singular value or is not
// finite, the cell centre when that is not finite either, then clamped to the cell.
V3 solve_cluster(const V3* pts, const V3* nrm, int m, uint32_t take, V3 lo, V3 hi, const SGrid& g,
                 Counters* ct) {
  int cnt = 0;
  for (int i = 0; i < m; ++i) cnt += (take >> i) & 1;
  const V3 centre = add(lo, mul(0.5f, v3(g.vx, g.vy, g.vz)));
  if (cnt == 0) return centre;
  V3 c = v3(0.f, 0.f, 0.f);
  for (int i = 0; i < m; ++i)
    if ((take >> i) & 1) c = add(c, pts[i]);
  c = mul(1.0f / (float)cnt, c);
  float B[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  float d[3] = {0.f, 0.f, 0.f};
  for (int i = 0; i < m; ++i) {
    if (!((take >> i) & 1)) continue;
    const V3 n = nrm[i];
    const float h = dot(n, sub(pts[i], c));
    B[0] += n.x * n.x;
    B[1] += n.x * n.y;
    B[2] += n.x * n.z;
    B[3] += n.y * n.x;
    B[4] += n.y * n.y;
    B[5] += n.y * n.z;
    B[6] += n.z * n.x;
    B[7] += n.z * n.y;
    B[8] += n.z * n.z;
    d[0] += n.x * h;
    d[1] += n.y * h;
    d[2] += n.z * h;
  }
  float y[3];
  bool singular = false;
  jacobi_pinv(B, d, 0.01f, y, &singular);
  V3 x = add(c, v3(y[0], y[1], y[2]));
  if (singular || !finite3(x)) {
    ++ct->fallbacks;
    x = c;
  }
  if (!finite3(x)) {
    ++ct->invalid;
    x = centre;
  }
  if (x.x < lo.x || x.x > hi.x || x.y < lo.y || x.y > hi.y || x.z < lo.z || x.z > hi.z) {
    ++ct->clamps;
    x.x = fmaxf(lo.x, fminf(hi.x, x.x));
    x.y = fmaxf(lo.y, fminf(hi.y, x.y));
    x.z = fmaxf(lo.z, fminf(hi.z, x.z));
  }
  return x;
}

// The hermite entries of one cell (dc_cuda_sparse.cu:488-563 and :682-758, the same code twice):
// per edge in kEdge order, a crossing (fa * fb <= 0) gives one entry at the linear zero with the
// normal of the chosen mode, a thin edge gives two entries at the |f|-weighted point with the two
// end gradients; at most 12 entries (a thin pair needs two free slots).
struct CellData {
  int m = 0;
  V3 p[12], n[12];
  float f[8];
  V3 c[8];
};

void cell_entries(const SGrid& g, int i, int j, int k, int mode, CellData* cd) {
  for (int q = 0; q < 8; ++q) {
    const int a = i + kCorner[q][0], b = j + kCorner[q][1], e = k + kCorner[q][2];
    cd->f[q] = g.at(a, b, e);
    cd->c[q] = v3(g.ox + (float)a * g.vx, g.oy + (float)b * g.vy, g.oz + (float)e * g.vz);
  }
  cd->m = 0;
  for (int e = 0; e < 12; ++e) {
    const int a = kEdge[e][0], b = kEdge[e][1];
    const float fa = cd->f[a], fb = cd->f[b];
    const int ia[3] = {i + kCorner[a][0], j + kCorner[a][1], k + kCorner[a][2]};
    const int ib[3] = {i + kCorner[b][0], j + kCorner[b][1], k + kCorner[b][2]};
    const bool crossing = fa * fb <= 0.0f;
    const bool thin = !crossing && thin_edge(g, fa, fb, ia, ib);
    if (!crossing && !thin) continue;
    float t = crossing ? fa / (fa - fb + 1e-12f) : fabsf(fa) / (fabsf(fa) + fabsf(fb) + 1e-12f);
    t = fmaxf(0.0f, fminf(1.0f, t));
    const V3 p = add(mul(1.0f - t, cd->c[a]), mul(t, cd->c[b]));
    if (thin) {
      V3 ga = sample_gradient(g, ia[0], ia[1], ia[2]);
      V3 gb = sample_gradient(g, ib[0], ib[1], ib[2]);
      const float la = sqrtf(dot(ga, ga)), lb = sqrtf(dot(gb, gb));
      if (la > 1e-6f) ga = mul(1.0f / la, ga);
      if (lb > 1e-6f) gb = mul(1.0f / lb, gb);
      if (cd->m < 11) {
        cd->p[cd->m] = p;
        cd->n[cd->m++] = ga;
        cd->p[cd->m] = p;
        cd->n[cd->m++] = gb;
      }
      continue;
    }
    V3 n;
    if (mode == 1) {  // EdgeGradient: the end gradients blended at t
      const V3 ga = sample_gradient(g, ia[0], ia[1], ia[2]);
      const V3 gb = sample_gradient(g, ib[0], ib[1], ib[2]);
      n = add(mul(1.0f - t, ga), mul(t, gb));
    } else {  // FiniteDifference; PrecomputedGradient without a gradient grid (:1310, :1336)
      n = trilinear_gradient(cd->f, cd->c, p);
    }
    const float ln = sqrtf(dot(n, n));
    if (ln > 1e-6f) n = mul(1.0f / ln, n);
    if (cd->m < 12) {
      cd->p[cd->m] = p;
      cd->n[cd->m++] = n;
    }
  }
}

// dc_cuda_sparse.cu:565-620: two vertices when the most opposed pair of entry normals has a dot
// under -0.3 and each side of the split (entries closer to one of the pair's normals) holds at
// least two entries. Returns 1 or 2; *take0 receives the first cluster's entries.
int split_cell(const CellData& cd, bool count_stats, Counters* ct, uint32_t* take0, int* ia,
               int* ib) {
  *take0 = (1u << cd.m) - 1u;
  *ia = *ib = -1;
  if (cd.m < 2) {
    if (count_stats) ++ct->one;
    return 1;
  }
  float lo = 1.0f;
  for (int a = 0; a < cd.m; ++a)
    for (int b = a + 1; b < cd.m; ++b) {
      const float t = dot(cd.n[a], cd.n[b]);
      if (t < lo) {
        lo = t;
        *ia = a;
        *ib = b;
      }
    }
  if (*ia < 0 || *ib < 0 || lo >= -0.3f) {
    if (count_stats) ++ct->one;
    return 1;
  }
  const V3 nA = cd.n[*ia], nB = cd.n[*ib];
  uint32_t t0 = 0;
  int c0 = 0, c1 = 0;
  for (int q = 0; q < cd.m; ++q) {
    if (dot(cd.n[q], nA) >= dot(cd.n[q], nB)) {
      t0 |= 1u << q;
      ++c0;
    } else {
      ++c1;
    }
  }
  if (c0 < 2 || c1 < 2) {
    if (count_stats) {
      ++ct->split_rejections;
      ++ct->one;
    }
    return 1;
  }
  if (count_stats) {
    ++ct->ambiguous;
    ++ct->multi;
    ++ct->two;
  }
  *take0 = t0;
  return 2;
}

struct SparseResult {
  std::vector<float> vertices;   // 3 per vertex
  std::vector<float> vnormals;   // 3 per vertex (multi-vertex form; zero for one-vertex cells)
  std::vector<int32_t> faces;    // 4 per quad
  std::vector<int64_t> active;   // active cells, ascending linear cell index
  std::vector<int32_t> vcount;   // vertices per active cell
  int64_t stats[16] = {};
};

// Vertex of cell `c` for a face whose edge gradient sum is gn (dc_cuda_sparse.cu:829-851).
int select_vertex(int64_t c, const std::vector<int32_t>& first, const std::vector<int8_t>& count,
                  const std::vector<float>& vn, V3 gn, bool mv) {
  const int s = first[c];
  if (s < 0) return -1;
  if (!mv) return s;
  const int n = count[c];
  if (n <= 0) return -1;
  if (n == 1) return s;
  const V3 n0 = v3(vn[3 * s], vn[3 * s + 1], vn[3 * s + 2]);
  const V3 n1 = v3(vn[3 * s + 3], vn[3 * s + 4], vn[3 * s + 5]);
  const float d0 = dot(gn, n0), d1 = dot(gn, n1);
  const int best = d0 >= d1 ? s : s + 1;
  if (fmaxf(d0, d1) < -0.5f) return -1;
  return best;
}

SparseResult* run_sparse(const SGrid& g, int B, int mode, bool mv) {
  SparseResult* R = new SparseResult();
  const int cx = g.nx - 1, cy = g.ny - 1, cz = g.nz - 1;
  const int64_t ncell = (int64_t)cx * cy * cz;
  const int nbx = (cx + B - 1) / B, nby = (cy + B - 1) / B, nbz = (cz + B - 1) / B;
  const int64_t nbrick = (int64_t)nbx * nby * nbz;
  const float near = 0.45f * g.vx;

  // dc_cuda_sparse.cu:271-308: a brick is active when its (B+1)^3 samples (clipped to the grid)
  // hold a negative and a positive value
  std::vector<uint8_t> brick(nbrick, 0);
  int64_t nab = 0;
  for (int bz = 0; bz < nbz; ++bz)
    for (int by = 0; by < nby; ++by)
      for (int bx = 0; bx < nbx; ++bx) {
        bool neg = false, pos = false;
        for (int k = bz * B; k < bz * B + B + 1 && k < g.nz; ++k)
          for (int j = by * B; j < by * B + B + 1 && j < g.ny; ++j)
            for (int i = bx * B; i < bx * B + B + 1 && i < g.nx; ++i) {
              const float v = g.at(i, j, k);
              neg = neg || v < 0.0f;
              pos = pos || v > 0.0f;
            }
        const int64_t b = bx + (int64_t)nbx * (by + (int64_t)nby * bz);
        brick[b] = neg && pos;
        nab += brick[b];
      }

  // :321-378: a cell of an active brick is active when its corners hold both signs or any corner
  // lies within 0.45 vx of the surface; the list ascends by linear cell index (:1277-1285)
  for (int k = 0; k < cz; ++k)
    for (int j = 0; j < cy; ++j)
      for (int i = 0; i < cx; ++i) {
        const int64_t b = i / B + (int64_t)nbx * (j / B + (int64_t)nby * (k / B));
        if (!brick[b]) continue;
        bool neg = false, pos = false, close = false;
        for (int q = 0; q < 8; ++q) {
          const float v = g.at(i + kCorner[q][0], j + kCorner[q][1], k + kCorner[q][2]);
          neg = neg || v < 0.0f;
          pos = pos || v > 0.0f;
          close = close || fabsf(v) <= near;
        }
        if ((neg && pos) || close) R->active.push_back(g.cell(i, j, k));
      }
  const int64_t na = (int64_t)R->active.size();

  Counters ct;
  // :445-620 (multi-vertex form only): vertices per active cell; their exclusive scan is each
  // cell's first vertex (:1316-1327)
  R->vcount.assign(na, 1);
  std::vector<CellData> cells(na);
  std::vector<int> ii(na), jj(na), kk(na);
  for (int64_t a = 0; a < na; ++a) {
    const int64_t c = R->active[a];
    ii[a] = (int)(c % cx);
    jj[a] = (int)((c / cx) % cy);
    kk[a] = (int)(c / ((int64_t)cx * cy));
    cell_entries(g, ii[a], jj[a], kk[a], mode, &cells[a]);
    if (mv) {
      uint32_t t0;
      int ia, ib;
      R->vcount[a] = split_cell(cells[a], true, &ct, &t0, &ia, &ib);
    }
  }
  std::vector<int32_t> first(ncell, -1);
  std::vector<int8_t> count(ncell, 0);
  int64_t nv = 0;
  for (int64_t a = 0; a < na; ++a) {
    first[R->active[a]] = (int32_t)nv;
    count[R->active[a]] = (int8_t)R->vcount[a];
    nv += R->vcount[a];
  }
  R->vertices.assign(3 * nv, 0.f);
  R->vnormals.assign(3 * nv, 0.f);

  // :622-827: one solve for a one-vertex cell, one per cluster for a split cell; a split cell's
  // vertex normals are its clusters' normalized mean normals (the pair's normals when zero)
  for (int64_t a = 0; a < na; ++a) {
    const CellData& cd = cells[a];
    const int s = first[R->active[a]];
    const V3 lo = cd.c[0], hi = cd.c[6];
    if (R->vcount[a] == 1 || cd.m < 2) {
      const V3 x = solve_cluster(cd.p, cd.n, cd.m, (1u << cd.m) - 1u, lo, hi, g, &ct);
      R->vertices[3 * s] = x.x;
      R->vertices[3 * s + 1] = x.y;
      R->vertices[3 * s + 2] = x.z;
      continue;
    }
    Counters scratch;
    uint32_t t0;
    int ia, ib;
    split_cell(cd, false, &scratch, &t0, &ia, &ib);
    const uint32_t all = (1u << cd.m) - 1u, t1 = all & ~t0;
    const V3 x0 = solve_cluster(cd.p, cd.n, cd.m, t0, lo, hi, g, &ct);
    const V3 x1 = solve_cluster(cd.p, cd.n, cd.m, t1, lo, hi, g, &ct);
    V3 s0 = v3(0.f, 0.f, 0.f), s1 = v3(0.f, 0.f, 0.f);
    for (int q = 0; q < cd.m; ++q) {
      if ((t0 >> q) & 1)
        s0 = add(s0, cd.n[q]);
      else
        s1 = add(s1, cd.n[q]);
    }
    const float l0 = sqrtf(dot(s0, s0)), l1 = sqrtf(dot(s1, s1));
    const V3 n0 = l0 > 1e-6f ? mul(1.0f / l0, s0) : cd.n[ia];
    const V3 n1 = l1 > 1e-6f ? mul(1.0f / l1, s1) : cd.n[ib];
    const V3 xs[2] = {x0, x1}, ns[2] = {n0, n1};
    for (int h = 0; h < 2; ++h) {
      R->vertices[3 * (s + h)] = xs[h].x;
      R->vertices[3 * (s + h) + 1] = xs[h].y;
      R->vertices[3 * (s + h) + 2] = xs[h].z;
      R->vnormals[3 * (s + h)] = ns[h].x;
      R->vnormals[3 * (s + h) + 1] = ns[h].y;
      R->vnormals[3 * (s + h) + 2] = ns[h].z;
    }
  }

  // :853-1102: per active cell its x, y and z edge from the cell's first corner; a strict sign
  // change gives a quad, a thin edge gives the quad and its reverse; the four cells around the
  // edge each give a vertex (a split cell the one whose normal best matches the edge gradients);
  // the quad turns to face along the summed end gradients.
  auto vert = [&](int v) { return v3(R->vertices[3 * v], R->vertices[3 * v + 1], R->vertices[3 * v + 2]); };
  for (int64_t a = 0; a < na; ++a) {
    const int i = ii[a], j = jj[a], k = kk[a];
    const size_t faces_before = R->faces.size();
    int skipped = 0;
    for (int axis = 0; axis < 3; ++axis) {
      int d1[3] = {0, 0, 0};
      d1[axis] = 1;
      // the four cells around the edge, in the reference's order (:871-874, :894-897, :917-920)
      int cc[4][3];
      if (axis == 0) {
        if (j < 1 || k < 1) continue;
        const int t[4][3] = {{i, j - 1, k - 1}, {i, j, k - 1}, {i, j, k}, {i, j - 1, k}};
        for (int q = 0; q < 4; ++q)
          for (int r = 0; r < 3; ++r) cc[q][r] = t[q][r];
      } else if (axis == 1) {
        if (i < 1 || k < 1) continue;
        const int t[4][3] = {{i - 1, j, k - 1}, {i, j, k - 1}, {i, j, k}, {i - 1, j, k}};
        for (int q = 0; q < 4; ++q)
          for (int r = 0; r < 3; ++r) cc[q][r] = t[q][r];
      } else {
        if (i < 1 || j < 1) continue;
        const int t[4][3] = {{i - 1, j - 1, k}, {i, j - 1, k}, {i, j, k}, {i - 1, j, k}};
        for (int q = 0; q < 4; ++q)
          for (int r = 0; r < 3; ++r) cc[q][r] = t[q][r];
      }
      const int ia[3] = {i, j, k}, ib[3] = {i + d1[0], j + d1[1], k + d1[2]};
      const float f0 = g.at(ia[0], ia[1], ia[2]), f1 = g.at(ib[0], ib[1], ib[2]);
      const bool crossing = f0 * f1 < 0.0f;
      const bool thin = !crossing && thin_edge(g, f0, f1, ia, ib);
      if (!crossing && !thin) continue;
      const V3 gn = add(sample_gradient(g, ia[0], ia[1], ia[2]),
                        sample_gradient(g, ib[0], ib[1], ib[2]));
      int v[4];
      bool ok = true;
      for (int q = 0; q < 4; ++q) {
        v[q] = select_vertex(g.cell(cc[q][0], cc[q][1], cc[q][2]), first, count, R->vnormals, gn, mv);
        ok = ok && v[q] >= 0;
      }
      if (!ok) {
        ++skipped;
        continue;
      }
      const V3 pa = vert(v[0]), pb = vert(v[1]), pc = vert(v[2]), pd = vert(v[3]);
      V3 qn = cross(sub(pb, pa), sub(pd, pa));
      if (dot(qn, qn) < 1e-12f) qn = cross(sub(pc, pa), sub(pd, pa));
      if (dot(gn, gn) > 0.0f && dot(qn, qn) > 0.0f && dot(qn, gn) < 0.0f) {
        const int t = v[1];
        v[1] = v[3];
        v[3] = t;
      }
      R->faces.insert(R->faces.end(), {v[0], v[1], v[2], v[3]});
      if (thin) R->faces.insert(R->faces.end(), {v[0], v[3], v[2], v[1]});
    }
    // the emission pass counts skipped edges only in cells that emit a face (:953, :1008-1010)
    if (mv && R->faces.size() > faces_before) ct.faces_skipped += skipped;
  }

  int64_t* st = R->stats;
  st[0] = ncell;
  st[1] = na;
  st[2] = nbrick;
  st[3] = nab;
  st[4] = nv;
  st[5] = (int64_t)R->faces.size() / 4;
  st[6] = ct.fallbacks;
  st[7] = ct.clamps;
  st[8] = ct.invalid;
  st[9] = ct.ambiguous;
  st[10] = ct.multi;
  st[11] = ct.one;
  st[12] = ct.two;
  st[13] = ct.split_rejections;
  st[14] = ct.bad_qef;
  st[15] = ct.faces_skipped;
  return R;
}

}  // namespace

extern "C" {

// Sparse extraction of a grid (values[i + nx*(j + ny*k)]); null for an invalid argument.
void* oracle_sparse_run(const float* values, int nx, int ny, int nz, float vx, float vy, float vz,
                        float ox, float oy, float oz, int brick_size, int normal_mode,
                        int multi_vertex) {
  if (!values || nx < 2 || ny < 2 || nz < 2 || brick_size < 1 || normal_mode < 0 ||
      normal_mode > 2)
    return nullptr;
  const SGrid g{values, nx, ny, nz, vx, vy, vz, ox, oy, oz};
  return run_sparse(g, brick_size, normal_mode, multi_vertex != 0);
}

// stats[16]: total cells, active cells, total bricks, active bricks, vertices, faces, QEF
// fallbacks, clamps, invalid, ambiguous, multi-vertex, one-vertex, two-vertex, split rejections,
// bad QEF, faces skipped
void oracle_sparse_stats(void* h, int64_t* stats) {
  const SparseResult* R = static_cast<SparseResult*>(h);
  for (int i = 0; i < 16; ++i) stats[i] = R->stats[i];
}

// vertices: 3 * stats[4] floats, vnormals: the same (may be null), faces: 4 * stats[5] ints,
// active: stats[1] int64 (may be null), vcount: stats[1] int32 (may be null)
void oracle_sparse_copy(void* h, float* vertices, float* vnormals, int32_t* faces, int64_t* active,
                        int32_t* vcount) {
  const SparseResult* R = static_cast<SparseResult*>(h);
  for (size_t i = 0; i < R->vertices.size(); ++i) vertices[i] = R->vertices[i];
  if (vnormals)
    for (size_t i = 0; i < R->vnormals.size(); ++i) vnormals[i] = R->vnormals[i];
  for (size_t i = 0; i < R->faces.size(); ++i) faces[i] = R->faces[i];
  if (active)
    for (size_t i = 0; i < R->active.size(); ++i) active[i] = R->active[i];
  if (vcount)
    for (size_t i = 0; i < R->vcount.size(); ++i) vcount[i] = R->vcount[i];
}

void oracle_sparse_free(void* h) { delete static_cast<SparseResult*>(h); }

// The 3x3 pseudo-inverse solve alone (pinning): returns 1 when every singular value was dropped.
int oracle_sparse_jacobi(const float* B, const float* d, float thr, float* y) {
  bool singular = false;
  jacobi_pinv(B, d, thr, y, &singular);
  return singular ? 1 : 0;
}

// One cell's hermite entries (pinning): returns m, pts and normals receive 3 * m floats.
int oracle_sparse_cell(const float* values, int nx, int ny, int nz, float vx, float vy, float vz,
                       float ox, float oy, float oz, int i, int j, int k, int normal_mode,
                       float* pts, float* normals) {
  const SGrid g{values, nx, ny, nz, vx, vy, vz, ox, oy, oz};
  CellData cd;
  cell_entries(g, i, j, k, normal_mode, &cd);
  for (int q = 0; q < cd.m; ++q) {
    pts[3 * q] = cd.p[q].x;
    pts[3 * q + 1] = cd.p[q].y;
    pts[3 * q + 2] = cd.p[q].z;
    normals[3 * q] = cd.n[q].x;
    normals[3 * q + 1] = cd.n[q].y;
    normals[3 * q + 2] = cd.n[q].z;
  }
  return cd.m;
}

}  // extern "C"
