// dc_cells_fast.hip — host launchers of the kernels whose results are held to tolerances, built
// with floating-point contraction (FMA) allowed: the FP32 mode's solve, its FP64 fallback cells,
// and the adjoint's per-cell kernels (both precisions).

#include "dc_cells_kernels.h"

hipError_t dc_launch_qef_fp32(const DcQefArgs& a, hipStream_t s) {
  const dim3 grid((unsigned)a.blocks), block(256);
  hipLaunchKernelGGL(k_qef<DC_PREC_FP32>, grid, block, 0, s, a.grids, a.n_grids, a.act_bounds,
                     a.act, a.alist, a.ctr_rw, a.capA, a.v64, a.v32, a.flags, a.fblist, a.ft, a.pw,
                     a.nrm, a.capF, a.qstate, a.rows);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_qef_fallback, dim3((unsigned)std::max<int64_t>(1, a.blocks * 2)), block,
                     0, s, a.grids, a.n_grids, a.act_bounds, a.act, a.alist, a.ctr_rw, a.capA,
                     a.fblist, a.v32, a.qstate64, a.cap64, a.ft, a.pw, a.nrm64, a.capF);
  return hipGetLastError();
}

namespace {
dim3 lane_grid(int64_t cells, int lanes) {
  return dim3((unsigned)std::max<int64_t>(1, (cells * lanes + 255) / 256));
}
}  // namespace

hipError_t dc_launch_bwd_phase(const DcBwdArgs& a, int phase, hipStream_t s, hipStream_t s64) {
  // the FP64 face normal table; FP32 mode's FP32 cells read its FP32 copy instead
  const double* n64 = reinterpret_cast<const double*>(a.nrm);
  const float* n32 = nullptr;
  const dim3 block(256);
  const int lanes = phase == 1 ? a.lanes1 : a.lanes2;  // for the active cells' kernel
  const dim3 grid = lane_grid(a.n_cells, lanes);
  // every launch takes the same arguments apart from the cell selection and the state
#define DC_BWD_ARGS(FLAGGED, LIST, STATE)                                                         \
  a.grids, a.n_grids, a.act_bounds, a.act, a.wprefix, a.alist, a.flags, FLAGGED, LIST, a.ctr,     \
      a.capA, a.grad_vertices, a.ft, a.pw, n64, n32, a.capF, STATE, a.ubar, a.ubar_e, a.rows,     \
      a.contrib
  // the active cells' kernel of phase P at `lanes` lanes per cell
#define DC_BWD_CELLS(P, T, S, GIVEN, FLAGGED, STATE)                                              \
  if (lanes == 16)                                                                                \
    hipLaunchKernelGGL((k_bwd_cell_lanes<P, T, S, 16>), grid, block, 0, s,                        \
                       DC_BWD_ARGS(FLAGGED, nullptr, STATE));                                     \
  else if (lanes == 4)                                                                            \
    hipLaunchKernelGGL((k_bwd_cell_lanes<P, T, S, 4>), grid, block, 0, s,                         \
                       DC_BWD_ARGS(FLAGGED, nullptr, STATE));                                     \
  else                                                                                            \
    hipLaunchKernelGGL((k_bwd_cell<P, T, S, GIVEN>), grid, block, 0, s,                           \
                       DC_BWD_ARGS(FLAGGED, nullptr, STATE));
  if (a.precision == DC_PREC_FP64) {
    if (phase == 1)  // solved again per cell: one thread per cell
      hipLaunchKernelGGL((k_bwd_cell<1, double, double, false>), lane_grid(a.n_cells, 1), block,
                         0, s, DC_BWD_ARGS(0, nullptr, nullptr));
    else {
      DC_BWD_CELLS(2, double, double, false, 0, nullptr)
    }
    return hipGetLastError();
  }
  // FP32 mode. Phase 1: the few FP64 cells first (on s64), at lanes64 lanes per cell: their long
  // chains start before the FP32 cells' blocks fill the device. They use their fallback solves'
  // states when all of them were kept, else solve every one again, one thread per cell (which
  // cells got a state depends on the order of the fallback list, so a mix would not be
  // reproducible).
  const dim3 fgrid = lane_grid(a.n_fallback, a.lanes64);
  const bool wide64 = a.lanes64 == 16;
  if (phase == 1) {  // s64: the FP64 cells (phase 2 runs on every cell below)
    if (a.fb_state && wide64)
      hipLaunchKernelGGL((k_bwd_cell_lanes<1, double, double, 16>), fgrid, block, 0, s64,
                         DC_BWD_ARGS(0, a.fblist, a.qstate64));
    else if (a.fb_state)
      hipLaunchKernelGGL((k_bwd_cell<1, double, double, true>), fgrid, block, 0, s64,
                         DC_BWD_ARGS(0, a.fblist, a.qstate64));
    else
      hipLaunchKernelGGL((k_bwd_cell<1, double, double, false>), lane_grid(a.n_fallback, 1),
                         block, 0, s64, DC_BWD_ARGS(0, a.fblist, nullptr));
  }
  {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  n64 = nullptr;
  n32 = a.nrm32;
  if (phase == 1) {
    DC_BWD_CELLS(1, float, float, true, 1, a.qstate)
  } else {
    // every cell: phase 2 takes each averaged row's face adjoint through the cell's own raw
    // normal, which does not depend on how well the cell's solve was conditioned
    DC_BWD_CELLS(2, float, float, false, 0, nullptr)
  }
#undef DC_BWD_CELLS
#undef DC_BWD_ARGS
  return hipGetLastError();
}

hipError_t dc_launch_bwd_face_adjoint(const DcBwdArgs& a, hipStream_t s) {
  const double* n64 = reinterpret_cast<const double*>(a.nrm);  // FP64 table in both modes
  const float* n32 = nullptr;
  if (a.precision == DC_PREC_FP64)
    hipLaunchKernelGGL(k_bwd_face_adjoint<double>, dim3((unsigned)a.face_blocks), dim3(256), 0, s,
                       a.ctr, a.capF, n64, n32, a.ubar);
  else
    hipLaunchKernelGGL(k_bwd_face_adjoint<float>, dim3((unsigned)a.face_blocks), dim3(256), 0, s,
                       a.ctr, a.capF, n64, n32, a.ubar);
  return hipGetLastError();
}
