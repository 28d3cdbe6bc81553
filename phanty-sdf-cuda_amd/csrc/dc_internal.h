// dc_internal.h — data layout shared by the host orchestration (dc_hip.cpp) and the CDNA4
// kernels (dc_kernels.hip). Not part of the public ABI.
//
// HBM layout for a batch of grids (a single extraction is a batch of one). All per-grid arrays
// are concatenated grid-major; offsets live in DcGridDev.
//
//   samples   float[nx*ny*nz] per grid, x fastest (caller-owned, DenseSdfGrid layout)
//   sg        sign planes: 16 bytes per sample pencil (x, y) and block of 64 sample layers,
//             entry (kb*ny + y)*64*Ws + x = {N, P} (uint64 each), bit t of N (P) set when the
//             sample (x, y, 64*kb + t) is < 0 (> 0); Ws = ceil(nx/64), Wks = ceil(nz/64)
//   act       uint64 active-cell plane, one word per 64 cells along x:
//             word (k*cy + j)*W + i/64, bit i%64; W = ceil(cx/64). Zero padding bits.
//   wprefix   uint32 per act word: exclusive count of active cells before it (global, so the
//             vertex id of cell (i,j,k) is wprefix[w] - vbase(grid) + popc(act[w] & below(bit))).
//   ft        uint64 face planes, one per axis a in {X,Y,Z}: bit k%64 of word
//             ((a*Wk + k/64)*cy + j)*cx + i says "quad on the axis-a edge at cell (i,j,k)";
//             Wk = ceil(cz/64). The k-bits of a pencil (a,i,j) are the face order the CPU
//             backend emits (contouring.cpp:609-611, 653-655, 693-695: k innermost).
//   pcount    uint32 faces per pencil (a,i,j), stored in face order (a*cx + i)*cy + j
//   pbase     uint32 exclusive scan of pcount over the batch = first face of the pencil
//   alist     uint64 active cells in vertex order: global act word * 64 + bit
//   v64/v32   vertices (FP64 for the plain solve, FP32 output), 3 per active cell
//   pw        uint32 per ft word: index of its first face (pbase + faces in earlier words)
//   flist     uint64 every face once as ft word * 64 + bit, in face-word (memory) order; its
//             face id is pw[word] + popc(ft[word] & below(bit))
//   scan      per scan (act words, pencils, face words): uint32 sum and uint64 base per tile of
//             1024 items
//   fnrm      per face, 4 of T (double plain, float FP32 mode): the normalized 4-cell averaged
//             hermite normal of the face edge and the length of the unnormalized sum
//   faces     int32 quads, 4 per face, grid-local vertex ids
// Adjoint:
//   ubar      float[4][3] per face: each member cell's adjoint of the averaged normal
//   ubar_e    float[12][3] per active cell: adjoint of an averaged normal that has no face (an
//             edge with a zero endpoint), per local edge
//   contrib   float[8] per active cell: its dL/dphi contributions to its own 8 corner samples
#pragma once
#include <stdint.h>

struct DcGridDev {
  const float* val;
  int32_t nx, ny, nz;
  int32_t cx, cy, cz;
  int32_t W, Wk;
  int32_t Ws;        // 64-sample words per sample row = ceil(nx/64)
  int32_t Wks;       // 64-layer blocks of sample layers = ceil(nz/64)
  float vx, vy, vz, ox, oy, oz;
  int64_t act_off;   // first act word
  int64_t sg_off;    // first sign-plane entry
  int64_t ft_off;    // first ft word
  int64_t pen_off;   // first pencil (pcount / pbase index)
  int64_t samp_off;  // first sample in a concatenated gradient buffer (backward)
  int64_t sr_off;    // first row group of the grid in the batch (adjoint gather; a group is
                     // gather_rows consecutive sample rows, y fastest, then z)
};

struct DcCounters {
  uint64_t n_active;      // total active cells (vertices), all grids
  uint64_t n_faces;       // total quads, all grids
  uint64_t fallback;      // FP32 mode: cells solved again in FP64
  uint64_t n_list;        // face list entries (equals n_faces)
  uint64_t n_reorient;    // FP32 mode: near-tie quads listed for FP64 orientation
  uint32_t nonfinite;     // a NaN or Inf sample was seen
  uint32_t error;         // internal failure bits: a list entry outside its grid (dc_decode.h)
};

enum DcPrecision { DC_PREC_FP64 = 0, DC_PREC_FP32 = 1 };
