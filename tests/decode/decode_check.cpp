// CPU check of the list-entry decoders (phanty-sdf-cuda_amd/csrc/dc_decode.h), compiled with g++
// by tests/test_decode_guards.py. For grids of several shapes: every cell's active-list entry
// decodes to that cell; entries with a padding bit (i >= cx), a word past the grid's plane or a
// word before it are rejected. Every face-plane bit decodes, and is accepted exactly when its edge
// is one the CPU backend's quad loops visit (contouring.cpp:607-729); bits past cz are rejected.
#include <cstdio>
#include <cstdlib>

#include "dc_decode.h"

using dcdecode::decode_cell;
using dcdecode::decode_face;

static int failures = 0;
#define CHECK(cond, ...)               \
  do {                                 \
    if (!(cond)) {                     \
      if (failures < 20) {             \
        std::printf(__VA_ARGS__);      \
        std::printf("\n");             \
      }                                \
      ++failures;                      \
    }                                  \
  } while (0)

static DcGridDev grid(int nx, int ny, int nz, int64_t act_off, int64_t ft_off) {
  DcGridDev g{};
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.cx = nx - 1;
  g.cy = ny - 1;
  g.cz = nz - 1;
  g.W = (g.cx + 63) / 64;
  g.Wk = (g.cz + 63) / 64;
  g.Ws = (nx + 63) / 64;
  g.Wks = (nz + 63) / 64;
  g.act_off = act_off;
  g.ft_off = ft_off;
  return g;
}

static void check_grid(int nx, int ny, int nz, int64_t act_off, int64_t ft_off) {
  const DcGridDev G = grid(nx, ny, nz, act_off, ft_off);
  const int64_t words = (int64_t)G.W * G.cy * G.cz;
  for (int k = 0; k < G.cz; ++k)
    for (int j = 0; j < G.cy; ++j)
      for (int i = 0; i < 64 * G.W; ++i) {
        const uint64_t w = (uint64_t)(act_off + ((int64_t)k * G.cy + j) * G.W + i / 64);
        int a, b, c;
        const bool ok = decode_cell(G, w * 64 + (uint64_t)(i % 64), a, b, c);
        if (i < G.cx)
          CHECK(ok && a == i && b == j && c == k, "cell %d %d %d of %dx%dx%d", i, j, k, nx, ny, nz);
        else
          CHECK(!ok, "padding bit i=%d accepted (%dx%dx%d)", i, nx, ny, nz);
      }
  int a, b, c;
  CHECK(!decode_cell(G, (uint64_t)(act_off + words) * 64, a, b, c), "word past the plane");
  if (act_off > 0) CHECK(!decode_cell(G, (uint64_t)(act_off - 1) * 64 + 5, a, b, c), "word before");
  const int64_t plane = (int64_t)G.cy * G.cx;
  for (int ax = 0; ax < 3; ++ax)
    for (int kw = 0; kw < G.Wk; ++kw)
      for (int j = 0; j < G.cy; ++j)
        for (int i = 0; i < G.cx; ++i)
          for (int bit = 0; bit < 64; ++bit) {
            const int k = 64 * kw + bit;
            const uint64_t w = (uint64_t)(ft_off + ((int64_t)ax * G.Wk + kw) * plane +
                                          (int64_t)j * G.cx + i);
            int fa, fi, fj, fk;
            const bool ok = decode_face(G, w * 64 + (uint64_t)bit, fa, fi, fj, fk);
            bool want = k < G.cz;
            if (ax == 0) want = want && j >= 1 && k >= 1;
            if (ax == 1) want = want && i >= 1 && k >= 1;
            if (ax == 2) want = want && i >= 1 && j >= 1;
            CHECK(ok == want, "face a=%d i=%d j=%d k=%d: %d, want %d", ax, i, j, k, ok, want);
            if (ok) CHECK(fa == ax && fi == i && fj == j && fk == k, "face coords a=%d", ax);
          }
  CHECK(!decode_face(G, (uint64_t)(ft_off + 3 * G.Wk * plane) * 64, a, b, c, c),
        "face word past the planes");
}

int main() {
  check_grid(2, 2, 2, 0, 0);
  check_grid(9, 8, 7, 0, 0);
  check_grid(66, 5, 70, 1000, 77);    // two words per row, two k-blocks, offsets of a batch
  check_grid(65, 3, 65, 12, 5);       // cx = 64: no padding bits; cz = 64: one full block
  check_grid(130, 4, 3, 3, 9);
  if (failures) {
    std::printf("%d failures\n", failures);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
