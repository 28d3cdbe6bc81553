// dc_kernels.h — host launchers for the kernels in dc_kernels.hip (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dc_internal.h"

struct DcMarkArgs {
  const DcGridDev* grids;
  int n_grids;
  const int64_t* tile_bounds;  // n_grids + 1
  int64_t n_tiles;
  int rows;             // sample rows per sign-pass wave tile: 4, 8, 16 or 32
  int waves_per_block;  // 1, 2 or 4
  uint4* sg;            // sign planes
  const int64_t* cw_bounds;  // n_grids + 1, cell-pass waves (W * cy per grid)
  int64_t n_cw;
  uint64_t* act;        // active words in vertex order [k][j][c]
  uint64_t* ft;
  uint32_t* pcount;
  DcCounters* ctr;
};

struct DcQefArgs {
  const DcGridDev* grids;
  int n_grids;
  const int64_t* act_bounds;  // n_grids + 1, in act words
  const uint64_t* act;
  const uint32_t* wprefix;
  const uint64_t* alist;
  const DcCounters* ctr;
  uint64_t capA;
  int64_t blocks;
  int precision;
  double* v64;
  float* v32;
  uint32_t* flags;   // per vertex, FP32 mode: 1 when the cell was solved in FP64
  uint32_t* fblist;  // FP32 mode: vertices to solve in FP64 (count in ctr->fallback)
  DcCounters* ctr_rw;
  const uint64_t* ft;
  const uint32_t* pw;
  const void* nrm;  // per-face averaged normals for the solve (FP64 table, or FP32 mode's copy)
  const double* nrm64;  // the FP64 table (FP32 mode's FP64 cells)
  uint64_t capF;
  float* qstate;    // FP32 mode: per-vertex solve state for the adjoint (24 floats), or null
  uint32_t* rows;   // FP32 mode: per-vertex row kinds for the adjoint (12), or null
  double* qstate64;  // FP32 mode: the FP64 cells' solve state by fallback list index (24 doubles)
  uint64_t cap64;    // entries of qstate64
};

// The three scans of one pass in shared launches (dc_launch_scan3): the active words (vertex ids
// and cell list), the pencil face counts (pencil bases) and the face words (face list).
struct DcScan3Args {
  const uint64_t* act;
  int64_t n_words;
  uint32_t* wprefix;
  uint64_t* alist;
  uint64_t capA;
  const uint32_t* pcount;
  int64_t n_pen;
  uint32_t* pbase;
  const uint64_t* ft;
  int64_t n_ft;
  uint64_t* flist;
  uint64_t capF;
  int64_t tiles_a, tiles_p, tiles_f;  // dc_scan_tiles of each
  uint32_t* tile_sum;   // tiles_a + tiles_p + tiles_f entries, in that order
  uint64_t* tile_base;
  DcCounters* ctr;      // n_active, n_faces, n_list
};

struct DcFaceNormalArgs {
  const DcGridDev* grids;
  int n_grids;
  const int64_t* pen_bounds;
  int64_t n_pen;
  const int64_t* ft_bounds;  // n_grids + 1, in face words
  int64_t n_ft;
  const uint64_t* ft;
  const uint32_t* pbase;
  uint32_t* pw;
  uint64_t* flist;
  uint32_t* tile_sum;   // face-list scan scratch, dc_scan_tiles(n_ft) entries each
  uint64_t* tile_base;
  DcCounters* ctr;
  uint64_t capF;
  int64_t face_blocks;  // grid-stride blocks of the per-face kernels
  int precision;
  void* nrm;     // FP64 table, 4 doubles per face
  float* nrm32;  // FP32 mode: its FP32 copy, 4 floats per face (null in plain mode)
  int list_done;        // the face list is already built (dc_launch_scan3)
};

struct DcEmitArgs {
  const DcGridDev* grids;
  int n_grids;
  const int64_t* ft_bounds;
  const uint64_t* flist;
  int64_t face_blocks;
  const uint64_t* ft;
  const uint32_t* pw;
  const uint64_t* act;
  const uint32_t* wprefix;
  const double* v64;
  const float* v32;
  int precision;
  DcCounters* ctr;  // read: the counts; written: error bits of out-of-grid face entries
  uint64_t capA;
  uint64_t capF;
  int32_t* faces;
  uint64_t* reorient;  // FP32 mode: capF entries for the near-tie quads' list entries, or null
  const double* nrm64;  // the FP64 face normal table (both modes)
};

// Adjoint gather: cell entries per pass in each wave's LDS list (8 bytes each; a multiple of 64).
constexpr int kGatherCap = 256;

struct DcBwdArgs {
  const DcGridDev* grids;
  int n_grids;
  const int64_t* act_bounds;
  const int64_t* sr_bounds;  // n_grids + 1, in gather row groups
  int64_t n_sr;
  int gather_rows;           // sample rows per gather group (1, 2, 4 or 8)
  int gather_lds_row;        // LDS floats per gathered row: 64 * min(16, max Ws) + 64
  const uint64_t* act;
  const uint32_t* wprefix;
  const uint64_t* alist;
  const uint32_t* flags;
  const uint32_t* fblist;
  int precision;
  const DcCounters* ctr;
  uint64_t capA;
  int64_t n_cells;          // active cells (host copy of the forward's count)
  int64_t n_fallback;       // FP32 mode's FP64 cells (host copy)
  int lanes1, lanes2;       // lanes per active cell in the phase 1 / phase 2 kernels: 1, 4 or 16
  int lanes64;              // lanes per FP64 cell of FP32 mode: 1 or 16
  int64_t gather_blocks;    // blocks of the gradient gather (0: one wave per row group)
  const float* grad_vertices;
  const uint64_t* ft;
  const uint32_t* pw;
  const void* nrm;      // FP64 face normal table
  const float* nrm32;   // FP32 mode: its FP32 copy (the FP32 cells' kernels read it)
  uint64_t capF;
  float* ubar;    // 4 slots x 3 floats per face (after the face pass: 3 doubles per face)
  float* ubar_e;  // 12 slots x 3 floats per active cell
  uint32_t* rows;  // 12 row kinds per active cell (FP32 mode: from the forward; else phase 1's)
  int64_t face_blocks;  // blocks of 256 over the faces
  float* contrib;  // 8 floats per active cell
  const float* qstate;  // FP32 mode: the forward's per-vertex solve state (24 floats), or null
  const double* qstate64;  // FP32 mode: the FP64 cells' solve state by fallback list index
  int fb_state;            // 1: qstate64 holds the state of every FP64 cell
  float* grad_values;
};

hipError_t dc_launch_fill_u64(uint64_t* p, int64_t n, uint64_t v, hipStream_t s);
hipError_t dc_launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s);
// Per-grid first vertex and first face of a batch (2 * n uint32, grid-major).
hipError_t dc_launch_grid_offsets(const int64_t* act_b, const int64_t* pen_b, int n,
                                  const uint32_t* wprefix, const uint32_t* pbase, uint32_t* out,
                                  hipStream_t s);
// Sign-pass tile rows: 4, 8, 16 or 32 sample rows of 64 samples per wave (k_mark), or
// kMarkWideRows: one 512-sample row segment per wave with 8 samples per lane (k_mark_wide; grids
// with nx % 8 == 0 and 16-byte aligned values).
constexpr int kMarkWideRows = 1;
int64_t dc_mark_tiles(const DcGridDev& g, int rows);  // sign-pass wave tiles of a grid
int64_t dc_cell_waves(const DcGridDev& g);            // cell-pass waves of a grid
// k_mark, the sign pass (sign planes); then dc_launch_mark_finish, the cell pass: the active plane
// in vertex order, the face planes over k and the pencil face counts.
hipError_t dc_launch_mark(const DcMarkArgs& a, hipStream_t s);
hipError_t dc_launch_mark_finish(const DcMarkArgs& a, hipStream_t s);
// Scans over n items in tiles of 1024: tile_sum (uint32) and tile_base (uint64) hold
// dc_scan_tiles(n) entries; *total receives the sum.
int64_t dc_scan_tiles(int64_t n);
hipError_t dc_launch_scan_act(const uint64_t* act, int64_t n_words, uint32_t* tile_sum,
                              uint64_t* tile_base, uint32_t* wprefix, uint64_t* alist,
                              uint64_t cap, uint64_t* total, hipStream_t s);
hipError_t dc_launch_scan_pen(const uint32_t* pcount, int64_t n_pen, uint32_t* tile_sum,
                              uint64_t* tile_base, uint32_t* pbase, uint64_t* total,
                              hipStream_t s);
// All three scans in three launches (reduce, tile scan, apply), each launch covering all three.
hipError_t dc_launch_scan3(const DcScan3Args& a, hipStream_t s);
hipError_t dc_launch_face_list(const uint64_t* ft, int64_t n_ft, uint32_t* tile_sum,
                               uint64_t* tile_base, uint64_t* flist, uint64_t cap,
                               uint64_t* total, hipStream_t s);
hipError_t dc_launch_face_normals(const DcFaceNormalArgs& a, hipStream_t s);
hipError_t dc_launch_qef(const DcQefArgs& a, hipStream_t s);
// Quad emission; in FP32 mode (a.reorient set) followed by the near-tie quads' FP64 orientation.
hipError_t dc_launch_emit(const DcEmitArgs& a, hipStream_t s);
// Adjoint: per-cell phase 1 or 2 (FP32 mode: the FP32-solved cells on s, the FP64 cells on s64,
// which may be another stream: the two sets are disjoint), then the gradient gather.
hipError_t dc_launch_bwd_phase(const DcBwdArgs& a, int phase, hipStream_t s, hipStream_t s64);
// Between the phases: every face's averaged-normal adjoint, once per face.
hipError_t dc_launch_bwd_face_adjoint(const DcBwdArgs& a, hipStream_t s);
hipError_t dc_launch_gather(const DcBwdArgs& a, hipStream_t s);
// triangulate_v1 of the last mesh: 2 nf x 3 int32 (goff: per-grid first vertex and face, n > 1).
hipError_t dc_launch_triangulate(const int32_t* faces, uint64_t nf, int n_grids,
                                 const uint32_t* goff, const double* v64, const float* v32,
                                 int use64, int32_t* tris, int64_t blocks, hipStream_t s);
// the contraction-enabled launchers (dc_cells_fast.hip)
hipError_t dc_launch_qef_fp32(const DcQefArgs& a, hipStream_t s);

