// dc_cells_fast.hip — host launchers of the kernels whose results are held to tolerances, built
// with floating-point contraction (FMA) allowed: the FP32 mode's face normals and solve, its FP64
// fallback cells, and the adjoint's per-cell kernels (both precisions).

#include "dc_cells_kernels.h"

hipError_t dc_launch_face_normals_fp32(const DcFaceNormalArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_face_normals<float>, dim3((unsigned)a.face_blocks), dim3(256), 0, s,
                     a.grids, a.n_grids, a.ft_bounds, a.flist, a.ft, a.pw, a.ctr, a.capF,
                     reinterpret_cast<float*>(a.nrm));
  return hipGetLastError();
}

hipError_t dc_launch_qef_fp32(const DcQefArgs& a, hipStream_t s) {
  const dim3 grid((unsigned)a.blocks), block(256);
  hipLaunchKernelGGL(k_qef<DC_PREC_FP32>, grid, block, 0, s, a.grids, a.n_grids, a.act_bounds,
                     a.act, a.alist, a.ctr_rw, a.capA, a.v64, a.v32, a.flags, a.fblist, a.ft, a.pw,
                     a.nrm, a.capF, a.qstate, a.rows);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_qef_fallback, dim3((unsigned)std::max<int64_t>(1, a.blocks * 2)), block,
                     0, s, a.grids, a.n_grids, a.act_bounds, a.act, a.alist, a.ctr_rw, a.capA,
                     a.fblist, a.v32, a.qstate64, a.cap64);
  return hipGetLastError();
}

hipError_t dc_launch_bwd_phase(const DcBwdArgs& a, int phase, hipStream_t s, hipStream_t s64) {
  const double* n64 = a.precision == DC_PREC_FP64 ? reinterpret_cast<const double*>(a.nrm) : nullptr;
  const float* n32 = a.precision == DC_PREC_FP64 ? nullptr : reinterpret_cast<const float*>(a.nrm);
  const dim3 grid((unsigned)a.cell_blocks), block(256);
  // every launch takes the same arguments apart from the cell selection and the state
#define DC_BWD_ARGS(FLAGGED, LIST, STATE)                                                         \
  a.grids, a.n_grids, a.act_bounds, a.act, a.wprefix, a.alist, a.flags, FLAGGED, LIST, a.ctr,     \
      a.capA, a.grad_vertices, a.ft, a.pw, n64, n32, a.capF, STATE, a.ubar, a.ubar_e, a.rows,     \
      a.contrib
  if (a.precision == DC_PREC_FP64) {
    if (phase == 1)
      hipLaunchKernelGGL((k_bwd_cell<1, double, double, false>), grid, block, 0, s,
                         DC_BWD_ARGS(0, nullptr, nullptr));
    else
      hipLaunchKernelGGL((k_bwd_cell<2, double, double, false>), grid, block, 0, s,
                         DC_BWD_ARGS(0, nullptr, nullptr));
    return hipGetLastError();
  }
  // FP32 mode. The few FP64 cells first (on s64): their long per-thread chains start before the
  // FP32 cells' blocks fill the device. Phase 1 of the FP64 cells uses their fallback solves'
  // states when all of them were kept, else solves every one again (which cells got a state
  // depends on the order of the fallback list, so a mix would not be reproducible).
  const dim3 fgrid((unsigned)a.fallback_blocks);
  if (phase == 1) {
    if (a.fb_state)
      hipLaunchKernelGGL((k_bwd_cell<1, double, double, true>), fgrid, block, 0, s64,
                         DC_BWD_ARGS(0, a.fblist, a.qstate64));
    else
      hipLaunchKernelGGL((k_bwd_cell<1, double, double, false>), fgrid, block, 0, s64,
                         DC_BWD_ARGS(0, a.fblist, nullptr));
  } else {
    hipLaunchKernelGGL((k_bwd_cell<2, double, float, false>), fgrid, block, 0, s64,
                       DC_BWD_ARGS(0, a.fblist, nullptr));
  }
  {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (phase == 1)
    hipLaunchKernelGGL((k_bwd_cell<1, float, float, true>), grid, block, 0, s,
                       DC_BWD_ARGS(1, nullptr, a.qstate));
  else
    hipLaunchKernelGGL((k_bwd_cell<2, float, float, false>), grid, block, 0, s,
                       DC_BWD_ARGS(1, nullptr, nullptr));
#undef DC_BWD_ARGS
  return hipGetLastError();
}

hipError_t dc_launch_bwd_face_adjoint(const DcBwdArgs& a, hipStream_t s) {
  const double* n64 = a.precision == DC_PREC_FP64 ? reinterpret_cast<const double*>(a.nrm) : nullptr;
  const float* n32 = a.precision == DC_PREC_FP64 ? nullptr : reinterpret_cast<const float*>(a.nrm);
  if (n64)
    hipLaunchKernelGGL(k_bwd_face_adjoint<double>, dim3((unsigned)a.face_blocks), dim3(256), 0, s,
                       a.ctr, a.capF, n64, n32, a.ubar);
  else
    hipLaunchKernelGGL(k_bwd_face_adjoint<float>, dim3((unsigned)a.face_blocks), dim3(256), 0, s,
                       a.ctr, a.capF, n64, n32, a.ubar);
  return hipGetLastError();
}
