// dc_internal.h — data layout shared by the host orchestration (dc_hip.cpp) and the CDNA4
// kernels (dc_kernels.hip). Not part of the public ABI.
//
// HBM layout for a batch of grids (a single extraction is a batch of one). All per-grid arrays
// are concatenated grid-major; offsets live in DcGridDev.
//
//   samples   float[nx*ny*nz] per grid, x fastest (caller-owned, DenseSdfGrid layout)
//   act       uint64 active-cell plane, one word per 64 cells along x:
//             word (k*cy + j)*W + i/64, bit i%64; W = ceil(cx/64). Zero padding bits.
//   wprefix   uint32 per act word: exclusive count of active cells before it (global, so the
//             vertex id of cell (i,j,k) is wprefix[w] - vbase(grid) + popc(act[w] & below(bit))).
//   ft        uint64 face planes, one per axis a in {X,Y,Z}: bit k%64 of word
//             ((a*Wk + k/64)*cy + j)*cx + i says "quad on the axis-a edge at cell (i,j,k)";
//             Wk = ceil(cz/64). The k-bits of a pencil (a,i,j) are the face order the CPU
//             backend emits (contouring.cpp:609-611, 653-655, 693-695: k innermost).
//   pcount    uint32 faces per pencil, stored (a*cy + j)*cx + i
//   pbase     uint32 exclusive scan of pcount in (grid, a, i, j) order = first face of the pencil
//   alist     uint64 active cells in vertex order: global act word * 64 + bit
//   v64/v32   vertices (FP64 for the plain solve, FP32 output), 3 per active cell
//   faces     int32 quads, 4 per face, grid-local vertex ids
//   contrib   float[64] per active cell: its dL/dphi contributions to the 4x4x4 samples around it
//   dil       uint64 candidate-sample plane: bit x of word (z*ny + y)*Ws + x/64 set when an
//             active cell's 4x4x4 stencil contains sample (x,y,z)
#pragma once
#include <stdint.h>

#define DC_MARK_ROWS_MAX 8

struct DcGridDev {
  const float* val;
  int32_t nx, ny, nz;
  int32_t cx, cy, cz;
  int32_t W, Wk;
  int32_t Ws;        // 64-bit words per sample row = ceil(nx/64) (candidate plane of the adjoint)
  int32_t pad_;
  float vx, vy, vz, ox, oy, oz;
  int64_t act_off;   // first act word
  int64_t ft_off;    // first ft word
  int64_t pen_off;   // first pencil (pcount / pbase index)
  int64_t samp_off;  // first sample in a concatenated gradient buffer (backward)
  int64_t dil_off;   // first word of the adjoint's candidate-sample plane
};

struct DcCounters {
  uint64_t n_active;      // total active cells (vertices), all grids
  uint64_t n_faces;       // total quads, all grids
  uint64_t fallback;      // FP32 mode: cells solved again in FP64
  uint32_t nonfinite;     // a NaN or Inf sample was seen
  uint32_t error;         // internal failure bits (scan wait limit)
};

// decoupled look-back tile status: [63:62] flag, [61:0] value
#define DC_SCAN_FLAG_AGG 1ull
#define DC_SCAN_FLAG_PREFIX 2ull
#define DC_SCAN_VALUE_MASK ((1ull << 62) - 1)
#define DC_ERR_SCAN_WAIT 1u

enum DcPrecision { DC_PREC_FP64 = 0, DC_PREC_FP32 = 1 };
