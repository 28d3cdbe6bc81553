// dc_standalone.cpp — definitions that the reference's dc_core provides inside the reference tree
// and that this repository's own builds (dc_cli) need beside HipDualContouringBackend. Built only
// against include/standalone/dc_backend.h; a build inside the reference tree leaves this file out
// (dc_core already defines these symbols).

#include <dc_backend.h>

#include <algorithm>

// src/cpp/dc_backend.cpp:8-49: clamped trilinear sample.
float DenseSdfGrid::sample(float x, float y, float z) const {
  const float lx = (x - ox) / vx, ly = (y - oy) / vy, lz = (z - oz) / vz;
  const int x0 = std::max(0, std::min((int)lx, nx - 2));
  const int y0 = std::max(0, std::min((int)ly, ny - 2));
  const int z0 = std::max(0, std::min((int)lz, nz - 2));
  const int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
  const float tx = std::max(0.0f, std::min(1.0f, lx - x0));
  const float ty = std::max(0.0f, std::min(1.0f, ly - y0));
  const float tz = std::max(0.0f, std::min(1.0f, lz - z0));
  auto at = [&](int i, int j, int k) { return values[i + nx * (j + ny * k)]; };
  const float c00 = at(x0, y0, z0) * (1.0f - tx) + at(x1, y0, z0) * tx;
  const float c01 = at(x0, y0, z1) * (1.0f - tx) + at(x1, y0, z1) * tx;
  const float c10 = at(x0, y1, z0) * (1.0f - tx) + at(x1, y1, z0) * tx;
  const float c11 = at(x0, y1, z1) * (1.0f - tx) + at(x1, y1, z1) * tx;
  const float c0 = c00 * (1.0f - ty) + c10 * ty;
  const float c1 = c01 * (1.0f - ty) + c11 * ty;
  return c0 * (1.0f - tz) + c1 * tz;
}
