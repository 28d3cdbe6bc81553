// m2s_oracle.cpp — CPU restatement of the reference's mesh-to-SDF producer,
// compute_mesh_sdf_device_cuda (src/cuda/mesh_to_sdf_cuda.cu:740-973), used ONLY as the parity
// checker of the GPU mesh-to-SDF pass.
//
// TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load this library. The product (phanty-sdf-cuda_amd/) never links, loads or calls it.
//
// Sequential in the reference's order, with its FP32 arithmetic (built with -ffp-contract=off):
//   bins        gx = max(1, nx/8) per axis, bin size (n*v)/g, triangle boxes grown by
//               2*max(v), clamped                                   mesh_to_sdf_cuda.cu:99-143
//   samples     p = o + i*v; closest point on each triangle of the 3x3x3 bins around p's bin
//               (the centre bin first, then the others in (z, y, x) order), ties within
//               1e-5 relative accumulate the weighted face normals (area, or the corner angle
//               when the closest point is a corner), sign from the summed normal, |d| beyond
//               2.5 bins -> FLT_MAX                                                    :226-360
//   flood       FLT_MAX samples joined to the exterior (the border, and positive samples) by
//               6-connected FLT_MAX paths become +10, the rest -10                     :362-428
//   relaxation  4 Jacobi passes: |f| > 1.5 min(v) takes min(|f|, |f_nb| + v_axis)       :430-456
//   voxelize    (option) voxels within 0.866 voxel diagonals of a triangle, dilated by one,
//               flooded from the border; enclosed voxels are solid; FLT_MAX samples take -10
//               in solid voxels, +10 elsewhere; no flood of the samples, no relaxation :458-679
// Two choices where the reference leaves the result open:
//   - The triangles of a bin are visited in ascending triangle order. The reference fills its bins
//     with atomics, so its order (which decides near-ties) varies from run to run; ascending order
//     is one of its outcomes.
//   - Floods run to their fixpoint. The reference's sample flood stops after 40 passes of an
//     in-place (unordered) update; where it converges, that is the same set.
// The corner angle is acos evaluated in FP64 and rounded to FP32 (the correctly rounded FP32
// arccos but for double rounding), within CUDA acosf's documented 2-ulp bound of the reference's.
// Triangle point distance: the closest-point regions of Ericson, "Real-Time Collision Detection"
// (2005), section 5.1.5, in the reference's operation order (mesh_to_sdf_cuda.cu:40-97).

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr float kFar = FLT_MAX;

struct V3 {
  float x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 scale(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// squared distance from p to triangle (a, b, c); the closest point goes to *q
float tri_dist_sq(V3 p, V3 a, V3 b, V3 c, V3* q) {
  const V3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
  const float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) {  // corner a
    *q = a;
    return dot(ap, ap);
  }
  const V3 bp = sub(p, b);
  const float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) {  // corner b
    *q = b;
    return dot(bp, bp);
  }
  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {  // edge ab
    const float t = d1 / (d1 - d3);
    *q = add(a, scale(t, ab));
    const V3 r = sub(p, *q);
    return dot(r, r);
  }
  const V3 cp = sub(p, c);
  const float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) {  // corner c
    *q = c;
    return dot(cp, cp);
  }
  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {  // edge ac
    const float t = d2 / (d2 - d6);
    *q = add(a, scale(t, ac));
    const V3 r = sub(p, *q);
    return dot(r, r);
  }
  const float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {  // edge bc
    const float t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    *q = add(b, scale(t, sub(c, b)));
    const V3 r = sub(p, *q);
    return dot(r, r);
  }
  const float inv = 1.0f / (va + vb + vc);  // interior
  const float s = vb * inv, t = vc * inv;
  *q = add(add(a, scale(s, ab)), scale(t, ac));
  const V3 r = sub(p, *q);
  return dot(r, r);
}

// floor(x) as a cell index clamped to [0, g - 1] (NaN and large values clamp instead of
// overflowing the integer conversion)
inline int clamp_index(float x, int g) {
  const float f = std::floor(x);
  if (!(f >= 0.0f)) return 0;
  if (f >= (float)(g - 1)) return g - 1;
  return (int)f;
}

// angle of the triangle's corner `k` (between its two edges), or -1 when an edge is too short
// (the area weight is kept then)
float corner_angle(V3 v0, V3 v1, V3 v2, int k) {
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
  const float l1 = std::sqrt(dot(e1, e1)), l2 = std::sqrt(dot(e2, e2));
  if (!(l1 > 1e-6f && l2 > 1e-6f)) return -1.0f;
  const float c = std::fmax(-1.0f, std::fmin(1.0f, dot(e1, e2) / (l1 * l2)));
  return (float)std::acos((double)c);
}

struct Input {
  const float* V;
  int64_t nv;
  const int32_t* F;
  int64_t nf;
  int nx, ny, nz;
  float ox, oy, oz, vx, vy, vz;
};

inline V3 vert(const Input& in, int32_t i) {
  return {in.V[3 * (int64_t)i], in.V[3 * (int64_t)i + 1], in.V[3 * (int64_t)i + 2]};
}

// Bins of the triangles whose grown box covers them, in ascending triangle order, and the
// per-triangle data the evaluation uses.
struct Bins {
  int gx, gy, gz;
  float bdx, bdy, bdz, min_dim;
  std::vector<std::vector<int32_t>> list;
};

void build_bins(const Input& in, Bins& B) {
  B.gx = std::max(1, in.nx / 8);
  B.gy = std::max(1, in.ny / 8);
  B.gz = std::max(1, in.nz / 8);
  B.bdx = ((float)in.nx * in.vx) / (float)B.gx;
  B.bdy = ((float)in.ny * in.vy) / (float)B.gy;
  B.bdz = ((float)in.nz * in.vz) / (float)B.gz;
  B.min_dim = std::fmin(B.bdx, std::fmin(B.bdy, B.bdz));
  const float margin = 2.0f * std::fmax(in.vx, std::fmax(in.vy, in.vz));
  B.list.assign((size_t)B.gx * B.gy * B.gz, {});
  for (int64_t f = 0; f < in.nf; ++f) {
    const V3 a = vert(in, in.F[3 * f]), b = vert(in, in.F[3 * f + 1]),
             c = vert(in, in.F[3 * f + 2]);
    const float lo[3] = {std::fmin(a.x, std::fmin(b.x, c.x)) - margin,
                         std::fmin(a.y, std::fmin(b.y, c.y)) - margin,
                         std::fmin(a.z, std::fmin(b.z, c.z)) - margin};
    const float hi[3] = {std::fmax(a.x, std::fmax(b.x, c.x)) + margin,
                         std::fmax(a.y, std::fmax(b.y, c.y)) + margin,
                         std::fmax(a.z, std::fmax(b.z, c.z)) + margin};
    const int x0 = clamp_index((lo[0] - in.ox) / B.bdx, B.gx),
              x1 = clamp_index((hi[0] - in.ox) / B.bdx, B.gx);
    const int y0 = clamp_index((lo[1] - in.oy) / B.bdy, B.gy),
              y1 = clamp_index((hi[1] - in.oy) / B.bdy, B.gy);
    const int z0 = clamp_index((lo[2] - in.oz) / B.bdz, B.gz),
              z1 = clamp_index((hi[2] - in.oz) / B.bdz, B.gz);
    for (int z = z0; z <= z1; ++z)
      for (int y = y0; y <= y1; ++y)
        for (int x = x0; x <= x1; ++x)
          B.list[(size_t)x + (size_t)B.gx * ((size_t)y + (size_t)B.gy * z)].push_back((int32_t)f);
  }
}

// The signed distance of one sample from the triangles of its 3x3x3 bins, or FLT_MAX. state
// (unless null) receives the final squared distance, closest point and summed normal.
float eval_sample(const Input& in, const Bins& B, int ix, int iy, int iz, float* state = nullptr) {
  const V3 p = {in.ox + (float)ix * in.vx, in.oy + (float)iy * in.vy, in.oz + (float)iz * in.vz};
  const int cx = clamp_index((p.x - in.ox) / B.bdx, B.gx);
  const int cy = clamp_index((p.y - in.oy) / B.bdy, B.gy);
  const int cz = clamp_index((p.z - in.oz) / B.bdz, B.gz);
  float best = kFar;
  V3 best_q = {0.0f, 0.0f, 0.0f}, best_n = {0.0f, 1.0f, 0.0f};
  const float eps_sq = 1e-6f * B.min_dim * B.min_dim;
  auto visit = [&](int bx, int by, int bz) {
    const std::vector<int32_t>& L =
        B.list[(size_t)bx + (size_t)B.gx * ((size_t)by + (size_t)B.gy * bz)];
    for (int32_t f : L) {
      const V3 v0 = vert(in, in.F[3 * (int64_t)f]), v1 = vert(in, in.F[3 * (int64_t)f + 1]),
               v2 = vert(in, in.F[3 * (int64_t)f + 2]);
      V3 q;
      const float d = tri_dist_sq(p, v0, v1, v2, &q);
      V3 n = cross(sub(v1, v0), sub(v2, v0));
      const float len = std::sqrt(dot(n, n));
      if (!(len > 1e-12f)) continue;  // degenerate: no normal, skipped
      n = {n.x / len, n.y / len, n.z / len};
      float w = len;
      const V3 r0 = sub(q, v0), r1 = sub(q, v1), r2 = sub(q, v2);
      int corner = -1;
      if (dot(r0, r0) < eps_sq)
        corner = 0;
      else if (dot(r1, r1) < eps_sq)
        corner = 1;
      else if (dot(r2, r2) < eps_sq)
        corner = 2;
      if (corner >= 0) {
        const float ang = corner_angle(v0, v1, v2, corner);
        if (ang >= 0.0f) w = ang;
      }
      const float tol = 1e-5f * (best + 1e-8f);
      if (d < best - tol) {
        best = d;
        best_q = q;
        best_n = scale(w, n);
      } else if (std::fabs(d - best) <= tol) {
        best_n = {best_n.x + w * n.x, best_n.y + w * n.y, best_n.z + w * n.z};
      }
    }
  };
  visit(cx, cy, cz);
  for (int bz = std::max(0, cz - 1); bz <= std::min(B.gz - 1, cz + 1); ++bz)
    for (int by = std::max(0, cy - 1); by <= std::min(B.gy - 1, cy + 1); ++by)
      for (int bx = std::max(0, cx - 1); bx <= std::min(B.gx - 1, cx + 1); ++bx)
        if (bx != cx || by != cy || bz != cz) visit(bx, by, bz);
  if (state) {
    const float st[7] = {best, best_q.x, best_q.y, best_q.z, best_n.x, best_n.y, best_n.z};
    std::memcpy(state, st, sizeof(st));
  }
  if (best == kFar) return kFar;
  float dist = std::sqrt(best);
  const float nl = std::sqrt(dot(best_n, best_n));
  if (nl > 1e-12f) best_n = {best_n.x / nl, best_n.y / nl, best_n.z / nl};
  if (dot(sub(p, best_q), best_n) < 0.0f) dist = -dist;
  const float lim = 2.5f * B.min_dim;
  if (dist > lim || dist < -lim) return kFar;
  return dist;
}

void evaluate(const Input& in, const Bins& B, float* out) {
  const int64_t nx = in.nx, ny = in.ny;
#pragma omp parallel for schedule(dynamic, 1)
  for (int iz = 0; iz < in.nz; ++iz)
    for (int iy = 0; iy < in.ny; ++iy)
      for (int ix = 0; ix < in.nx; ++ix)
        out[ix + nx * (iy + ny * (int64_t)iz)] = eval_sample(in, B, ix, iy, iz);
}

// Marks every cell reachable from the seeds through passable cells (6-connectivity). ext holds
// the seeds on entry and the reached set on return.
void flood(int nx, int ny, int nz, const std::vector<uint8_t>& pass, std::vector<uint8_t>& ext) {
  const int64_t sx = 1, sy = nx, sz = (int64_t)nx * ny;
  std::vector<int64_t> stack;
  for (int64_t i = 0; i < (int64_t)ext.size(); ++i)
    if (ext[i]) stack.push_back(i);
  while (!stack.empty()) {
    const int64_t i = stack.back();
    stack.pop_back();
    const int64_t x = i % nx, y = (i / nx) % ny, z = i / sz;
    const int64_t nb[6] = {x > 0 ? i - sx : -1,      x < nx - 1 ? i + sx : -1,
                           y > 0 ? i - sy : -1,      y < ny - 1 ? i + sy : -1,
                           z > 0 ? i - sz : -1,      z < nz - 1 ? i + sz : -1};
    for (int64_t j : nb)
      if (j >= 0 && !ext[j] && pass[j]) {
        ext[j] = 1;
        stack.push_back(j);
      }
  }
}

inline bool on_border(int x, int y, int z, int nx, int ny, int nz) {
  return x == 0 || x == nx - 1 || y == 0 || y == ny - 1 || z == 0 || z == nz - 1;
}

// FLT_MAX samples: +10 where joined to the exterior, -10 in enclosed cavities
void resolve_far(const Input& in, float* out) {
  const int nx = in.nx, ny = in.ny, nz = in.nz;
  const int64_t n = (int64_t)nx * ny * nz;
  std::vector<uint8_t> pass(n), ext(n);
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const int64_t i = x + (int64_t)nx * (y + (int64_t)ny * z);
        const float f = out[i];
        pass[i] = f == kFar;
        // the border's far and positive samples, and every other positive sample
        ext[i] = f > 0.0f && (on_border(x, y, z, nx, ny, nz) || f != kFar);
      }
  flood(nx, ny, nz, pass, ext);
  for (int64_t i = 0; i < n; ++i)
    if (out[i] == kFar) out[i] = ext[i] ? 10.0f : -10.0f;
}

void relax(const Input& in, float* out) {
  const int64_t nx = in.nx, ny = in.ny, nz = in.nz, n = nx * ny * nz;
  const float keep = std::fmin(in.vx, std::fmin(in.vy, in.vz)) * 1.5f;
  std::vector<float> tmp(n);
  float* src = out;
  float* dst = tmp.data();
  for (int pass = 0; pass < 4; ++pass) {
#pragma omp parallel for
    for (int64_t z = 0; z < nz; ++z)
      for (int64_t y = 0; y < ny; ++y)
        for (int64_t x = 0; x < nx; ++x) {
          const int64_t i = x + nx * (y + ny * z);
          const float f = src[i];
          if (std::fabs(f) <= keep) {
            dst[i] = f;
            continue;
          }
          float m = std::fabs(f);
          if (x > 0) m = std::fmin(m, std::fabs(src[i - 1]) + in.vx);
          if (x < nx - 1) m = std::fmin(m, std::fabs(src[i + 1]) + in.vx);
          if (y > 0) m = std::fmin(m, std::fabs(src[i - nx]) + in.vy);
          if (y < ny - 1) m = std::fmin(m, std::fabs(src[i + nx]) + in.vy);
          if (z > 0) m = std::fmin(m, std::fabs(src[i - nx * ny]) + in.vz);
          if (z < nz - 1) m = std::fmin(m, std::fabs(src[i + nx * ny]) + in.vz);
          dst[i] = (f < 0.0f ? -1.0f : 1.0f) * m;
        }
    std::swap(src, dst);
  }
  // four passes end in `out`
}

// Solid voxels (1) of the voxelized mesh: shell voxels within 0.866 voxel diagonals of a
// triangle, dilated by one voxel, plus every voxel the border's flood does not reach.
void voxelize(const Input& in, int vres, float vox_x, float vox_y, float vox_z,
              std::vector<uint8_t>& solid) {
  const int64_t n = (int64_t)vres * vres * vres;
  std::vector<uint8_t> shell(n, 0);
  const float reach = std::fmax(vox_x, std::fmax(vox_y, vox_z)) * 0.8660254f;
  const float reach_sq = reach * reach;
  for (int64_t f = 0; f < in.nf; ++f) {
    const V3 a = vert(in, in.F[3 * f]), b = vert(in, in.F[3 * f + 1]),
             c = vert(in, in.F[3 * f + 2]);
    const int x0 = clamp_index((std::fmin(a.x, std::fmin(b.x, c.x)) - in.ox) / vox_x, vres);
    const int x1 = clamp_index((std::fmax(a.x, std::fmax(b.x, c.x)) - in.ox) / vox_x, vres);
    const int y0 = clamp_index((std::fmin(a.y, std::fmin(b.y, c.y)) - in.oy) / vox_y, vres);
    const int y1 = clamp_index((std::fmax(a.y, std::fmax(b.y, c.y)) - in.oy) / vox_y, vres);
    const int z0 = clamp_index((std::fmin(a.z, std::fmin(b.z, c.z)) - in.oz) / vox_z, vres);
    const int z1 = clamp_index((std::fmax(a.z, std::fmax(b.z, c.z)) - in.oz) / vox_z, vres);
    for (int z = z0; z <= z1; ++z)
      for (int y = y0; y <= y1; ++y)
        for (int x = x0; x <= x1; ++x) {
          const V3 ctr = {in.ox + (float)x * vox_x, in.oy + (float)y * vox_y,
                          in.oz + (float)z * vox_z};
          V3 q;
          if (tri_dist_sq(ctr, a, b, c, &q) <= reach_sq)
            shell[x + (int64_t)vres * (y + (int64_t)vres * z)] = 1;
        }
  }
  std::vector<uint8_t> grown(n, 0);
#pragma omp parallel for
  for (int z = 0; z < vres; ++z)
    for (int y = 0; y < vres; ++y)
      for (int x = 0; x < vres; ++x) {
        uint8_t any = 0;
        for (int dz = std::max(0, z - 1); dz <= std::min(vres - 1, z + 1) && !any; ++dz)
          for (int dy = std::max(0, y - 1); dy <= std::min(vres - 1, y + 1) && !any; ++dy)
            for (int dx = std::max(0, x - 1); dx <= std::min(vres - 1, x + 1); ++dx)
              if (shell[dx + (int64_t)vres * (dy + (int64_t)vres * dz)]) {
                any = 1;
                break;
              }
        grown[x + (int64_t)vres * (y + (int64_t)vres * z)] = any;
      }
  std::vector<uint8_t> pass(n), ext(n);
  for (int z = 0; z < vres; ++z)
    for (int y = 0; y < vres; ++y)
      for (int x = 0; x < vres; ++x) {
        const int64_t i = x + (int64_t)vres * (y + (int64_t)vres * z);
        pass[i] = !grown[i];
        ext[i] = !grown[i] && on_border(x, y, z, vres, vres, vres);
      }
  flood(vres, vres, vres, pass, ext);
  solid.resize(n);
  for (int64_t i = 0; i < n; ++i) solid[i] = !ext[i];
}

struct VoxelGrid {
  int vres;
  float vx, vy, vz;
};

VoxelGrid voxel_grid(const Input& in, int voxel_res) {
  VoxelGrid G;
  G.vres = std::max(32, std::min(2048, voxel_res));
  G.vx = ((float)in.nx * in.vx) / (float)G.vres;
  G.vy = ((float)in.ny * in.vy) / (float)G.vres;
  G.vz = ((float)in.nz * in.vz) / (float)G.vres;
  return G;
}

}  // namespace

extern "C" {

// The reference's compute_mesh_sdf_device_cuda values for this mesh and grid (out: nx*ny*nz
// floats). Returns 0, or -1 when a triangle names a vertex outside [0, nv).
int oracle_mesh_to_sdf(const float* V, int64_t nv, const int32_t* F, int64_t nf, int nx, int ny,
                       int nz, float ox, float oy, float oz, float vx, float vy, float vz,
                       int voxelize_first, int voxel_res, float* out) {
  for (int64_t i = 0; i < 3 * nf; ++i)
    if (F[i] < 0 || F[i] >= nv) return -1;
  const Input in{V, nv, F, nf, nx, ny, nz, ox, oy, oz, vx, vy, vz};
  Bins B;
  build_bins(in, B);
  evaluate(in, B, out);
  if (!voxelize_first) {
    resolve_far(in, out);
    relax(in, out);
    return 0;
  }
  const VoxelGrid G = voxel_grid(in, voxel_res);
  std::vector<uint8_t> solid;
  voxelize(in, G.vres, G.vx, G.vy, G.vz, solid);
  const int64_t n = (int64_t)nx * ny * nz;
  for (int64_t i = 0; i < n; ++i) {
    if (out[i] != kFar) continue;
    const int64_t x = i % nx, y = (i / nx) % ny, z = i / ((int64_t)nx * ny);
    const float px = ox + (float)x * vx, py = oy + (float)y * vy, pz = oz + (float)z * vz;
    const int bx = clamp_index((px - ox) / G.vx, G.vres);
    const int by = clamp_index((py - oy) / G.vy, G.vres);
    const int bz = clamp_index((pz - oz) / G.vz, G.vres);
    out[i] = solid[bx + (int64_t)G.vres * (by + (int64_t)G.vres * bz)] ? -10.0f : 10.0f;
  }
  return 0;
}

// The evaluation alone (FLT_MAX where no triangle is within 2.5 bins), before the flood.
int oracle_mesh_to_sdf_raw(const float* V, int64_t nv, const int32_t* F, int64_t nf, int nx,
                           int ny, int nz, float ox, float oy, float oz, float vx, float vy,
                           float vz, float* out) {
  for (int64_t i = 0; i < 3 * nf; ++i)
    if (F[i] < 0 || F[i] >= nv) return -1;
  const Input in{V, nv, F, nf, nx, ny, nz, ox, oy, oz, vx, vy, vz};
  Bins B;
  build_bins(in, B);
  evaluate(in, B, out);
  return 0;
}

// The solid voxels (vres^3 bytes, x fastest) of the voxelize option; returns vres.
int oracle_m2s_voxels(const float* V, int64_t nv, const int32_t* F, int64_t nf, int nx, int ny,
                      int nz, float ox, float oy, float oz, float vx, float vy, float vz,
                      int voxel_res, uint8_t* solid_out) {
  const Input in{V, nv, F, nf, nx, ny, nz, ox, oy, oz, vx, vy, vz};
  const VoxelGrid G = voxel_grid(in, voxel_res);
  if (!solid_out) return G.vres;
  std::vector<uint8_t> solid;
  voxelize(in, G.vres, G.vx, G.vy, G.vz, solid);
  std::memcpy(solid_out, solid.data(), solid.size());
  return G.vres;
}

// One sample's evaluation state: out[0] squared distance, out[1..3] closest point, out[4..6] the
// summed weighted normal (before normalisation).
void oracle_m2s_sample_state(const float* V, int64_t nv, const int32_t* F, int64_t nf, int nx,
                             int ny, int nz, float ox, float oy, float oz, float vx, float vy,
                             float vz, int ix, int iy, int iz, float* out) {
  const Input in{V, nv, F, nf, nx, ny, nz, ox, oy, oz, vx, vy, vz};
  Bins B;
  build_bins(in, B);
  eval_sample(in, B, ix, iy, iz, out);
}

// Squared distance from p to triangle (a, b, c) and the closest point (3 floats each).
float oracle_point_triangle(const float* p, const float* a, const float* b, const float* c,
                            float* closest) {
  V3 q;
  const float d = tri_dist_sq({p[0], p[1], p[2]}, {a[0], a[1], a[2]}, {b[0], b[1], b[2]},
                              {c[0], c[1], c[2]}, &q);
  closest[0] = q.x;
  closest[1] = q.y;
  closest[2] = q.z;
  return d;
}

}  // extern "C"
