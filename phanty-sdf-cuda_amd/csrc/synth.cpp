// synth.cpp — synthetic dense SDF grids for the benchmarks and the CLI (host code, OpenMP).
//
// Primitive formulas follow the reference's analytic SDF factories:
//   sphere          |x - c| - r                                     src/cpp/sdf.cpp:101-116
//   rotated box     |R^T (x - c)| - b, outside + inside distance    src/cpp/sdf.cpp:46-78
//   capped cylinder Y axis, full height he, in the rotated frame     src/cpp/sdf.cpp:119-155
// and the dc_cli generators (FP32 host loops)                        src/cpp/dc_cli.cpp:49-118.
//
// Unit-cube grids: n samples per axis, x_i = i / (n - 1) evaluated in FP64, value rounded to FP32.

#include <algorithm>
#include <cmath>
#include <cstdint>

namespace {

struct Prim {
  int type;  // 0 sphere (p0 = r), 1 box (p0..p2 half extents), 2 capped cylinder (p0 = r, p1 = he)
  double c[3];
  double R[9];  // row-major; local = R^T (x - c)
  double p[3];
};

inline double prim_eval(const Prim& q, const double x[3]) {
  const double t[3] = {x[0] - q.c[0], x[1] - q.c[1], x[2] - q.c[2]};
  if (q.type == 0) return std::sqrt((t[0] * t[0] + t[1] * t[1]) + t[2] * t[2]) - q.p[0];
  double l[3];
  for (int a = 0; a < 3; ++a) l[a] = (q.R[0 + a] * t[0] + q.R[3 + a] * t[1]) + q.R[6 + a] * t[2];
  if (q.type == 1) {
    const double qq[3] = {std::abs(l[0]) - q.p[0], std::abs(l[1]) - q.p[1], std::abs(l[2]) - q.p[2]};
    const double qp[3] = {std::max(qq[0], 0.0), std::max(qq[1], 0.0), std::max(qq[2], 0.0)};
    const double outside = std::sqrt((qp[0] * qp[0] + qp[1] * qp[1]) + qp[2] * qp[2]);
    const double inside = std::min(std::max(qq[0], std::max(qq[1], qq[2])), 0.0);
    return outside + inside;
  }
  const double rr = std::sqrt(l[0] * l[0] + l[2] * l[2]);
  const double half = q.p[1] / 2.0;
  const double g = std::max(rr - q.p[0], std::abs(l[1]) - half);
  const double h0 = std::max(rr - q.p[0], 0.0), h1 = std::max(std::abs(l[1]) - half, 0.0);
  const double f = std::sqrt(h0 * h0 + h1 * h1);
  return (g <= 0.0) ? g : f;
}

// Polynomial smooth minimum (Quilez): min(a,b) - max(k - |a-b|, 0)^2 / (4k).
inline double smin(double a, double b, double k) {
  const double h = std::max(k - std::abs(a - b), 0.0);
  return std::min(a, b) - h * h / (4.0 * k);
}

}  // namespace

extern "C" {

// prims: nprim records of 16 doubles: {type, cx, cy, cz, R[9], p0, p1, p2}.
// ops[q] for q >= 1 combines the running value a with primitive b:
//   0: min(a, b)    1: max(a, -b)    2: smin(a, b; k)
// clampv > 0 clamps the result to [-clampv, clampv]. out has n^3 floats, x fastest.
int dc_synth_csg(int n, int nprim, const double* prims, const int* ops, double k, double clampv,
                 float* out) {
  if (n < 2 || nprim < 1) return 1;
  Prim P[256];
  if (nprim > 256) return 1;
  for (int q = 0; q < nprim; ++q) {
    const double* r = prims + 16 * q;
    P[q].type = (int)r[0];
    for (int a = 0; a < 3; ++a) P[q].c[a] = r[1 + a];
    for (int a = 0; a < 9; ++a) P[q].R[a] = r[4 + a];
    for (int a = 0; a < 3; ++a) P[q].p[a] = r[13 + a];
  }
  const double h = 1.0 / (double)(n - 1);
#pragma omp parallel for schedule(static)
  for (int kk = 0; kk < n; ++kk) {
    for (int j = 0; j < n; ++j) {
      float* row = out + (int64_t)n * (j + (int64_t)n * kk);
      for (int i = 0; i < n; ++i) {
        const double x[3] = {i * h, j * h, kk * h};
        double v = prim_eval(P[0], x);
        for (int q = 1; q < nprim; ++q) {
          const double b = prim_eval(P[q], x);
          if (ops[q] == 0)
            v = std::min(v, b);
          else if (ops[q] == 1)
            v = std::max(v, -b);
          else
            v = smin(v, b, k);
        }
        if (clampv > 0.0) v = std::min(std::max(v, -clampv), clampv);
        row[i] = (float)v;
      }
    }
  }
  return 0;
}

// dc_cli.cpp:49-118 generators on [-1,1]^3 with spacing 2/(res-1), in FP32 like the reference.
// kind: 0 sphere (r = 0.5), 1 box (half extent 0.5), 2 plane (value z).
int dc_synth_cli(int kind, int res, float* out) {
  if (res < 2) return 1;
  const float v = 2.0f / (res - 1);
  const float o = -1.0f;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < res; ++k) {
    const float z = o + k * v;
    for (int j = 0; j < res; ++j) {
      const float y = o + j * v;
      for (int i = 0; i < res; ++i) {
        const float x = o + i * v;
        float val;
        if (kind == 0) {
          val = std::sqrt(x * x + y * y + z * z) - 0.5f;
        } else if (kind == 1) {
          const float dx = std::abs(x) - 0.5f, dy = std::abs(y) - 0.5f, dz = std::abs(z) - 0.5f;
          val = std::max(dx, std::max(dy, dz));
        } else {
          val = z;
        }
        out[i + (int64_t)res * (j + (int64_t)res * k)] = val;
      }
    }
  }
  return 0;
}

}  // extern "C"
