// dc_decode.h — list entries to grid coordinates, with the bounds every consumer relies on.
//
// The per-cell kernels read a cell's 8 corner samples and the per-face kernels the 4 cells around
// a grid edge; both take their coordinates from list entries built by the marking pass and the
// scans. A stray bit in the active plane or a face plane (a bug upstream) would otherwise become
// a corner load past the grid. Each decoder returns false for an entry outside its grid; the
// kernels then skip the entry and set a bit in DcCounters::error, which the host reports as
// DC_HIP_ERR_INTERNAL. Host and device code (the CPU tests compile this header with g++).
#pragma once
#include <stdint.h>

#include "dc_internal.h"

#if defined(__HIPCC__)
#define DC_HD __host__ __device__
#else
#define DC_HD
#endif

namespace dcdecode {

// DcCounters::error bits
constexpr uint32_t kErrCellRange = 1u;  // an active-list entry named a cell outside its grid
constexpr uint32_t kErrFaceRange = 2u;  // a face-list entry named an edge that emits no quad

// Active-list entry (global act word * 64 + bit) -> cell (i, j, k) of grid G. The act word of
// cell (i, j, k) is act_off + (k*cy + j)*W + i/64 (vertex order). False unless the word lies in
// G's plane and i < cx (the padding bits of a row's last word name no cell).
DC_HD inline bool decode_cell(const DcGridDev& G, uint64_t entry, int& i, int& j, int& k) {
  const int64_t lw = (int64_t)(entry >> 6) - G.act_off;
  const int64_t row = lw / G.W;
  const int c = (int)(lw - row * G.W);
  i = 64 * c + (int)(entry & 63);
  j = (int)(row % G.cy);
  k = (int)(row / G.cy);
  return lw >= 0 && row < (int64_t)G.cy * G.cz && i < G.cx;
}

// Does the axis-a grid edge with lower endpoint (i, j, k) emit a quad when its sign changes
// (contouring.cpp:607-729)? x edges i < cx, 1 <= j < cy, 1 <= k < cz; y edges 1 <= i < cx,
// j < cy, 1 <= k < cz; z edges 1 <= i < cx, 1 <= j < cy, k < cz. Their four cells (one step back
// across the edge on each of the other two axes) all lie in the grid.
DC_HD inline bool face_edge_ok(const DcGridDev& G, int a, int i, int j, int k) {
  if (a < 0 || a > 2 || i >= G.cx || j >= G.cy || k >= G.cz) return false;
  const int lo_i = a == 0 ? 0 : 1, lo_j = a == 1 ? 0 : 1, lo_k = a == 2 ? 0 : 1;
  return i >= lo_i && j >= lo_j && k >= lo_k;
}

// Face-plane bit (global face word * 64 + bit) -> the quad's grid edge: axis a and lower endpoint
// (i, j, k). The face word of axis a, k-block kw and pencil (i, j) is
// ft_off + ((a*Wk + kw)*cy + j)*cx + i, bit k % 64. True only for an edge in G's planes that can
// emit a quad (face_edge_ok).
DC_HD inline bool decode_face(const DcGridDev& G, uint64_t entry, int& a, int& i, int& j, int& k) {
  const int64_t plane = (int64_t)G.cy * G.cx;
  const int64_t local = (int64_t)(entry >> 6) - G.ft_off;
  const int64_t aw = local / plane;  // a * Wk + kw
  const int64_t rem = local - aw * plane;
  a = (int)(aw / G.Wk);
  const int kw = (int)(aw - (int64_t)a * G.Wk);
  j = (int)(rem / G.cx);
  i = (int)(rem - (int64_t)j * G.cx);
  k = 64 * kw + (int)(entry & 63);
  return local >= 0 && face_edge_ok(G, a, i, j, k);
}

}  // namespace dcdecode
