// dc_cells_fast.hip — host launchers of the kernels whose results are held to tolerances, built
// with floating-point contraction (FMA) allowed: the FP32 mode's face normals and solve, its FP64
// fallback cells, and the adjoint's per-cell kernels (both precisions).

#include <type_traits>

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
                     a.fblist, a.v32);
  return hipGetLastError();
}

hipError_t dc_launch_bwd_phase(const DcBwdArgs& a, int phase, hipStream_t s, hipStream_t s64) {
  const double* n64 = a.precision == DC_PREC_FP64 ? reinterpret_cast<const double*>(a.nrm) : nullptr;
  const float* n32 = a.precision == DC_PREC_FP64 ? nullptr : reinterpret_cast<const float*>(a.nrm);
  const dim3 grid((unsigned)a.cell_blocks), block(256);
  const dim3 fgrid((unsigned)a.fallback_blocks);
// Phase 2 reads the faces' adjoints as double in plain mode and as float in FP32 mode (S); phase 1
// does not read them, so its FP64 instance serves both modes.
#define DC_BWD(PHASE)                                                                              \
  if (a.precision == DC_PREC_FP64) {                                                               \
    hipLaunchKernelGGL((k_bwd_cell<PHASE, double, double>), grid, block, 0, s, a.grids,            \
                       a.n_grids, a.act_bounds, a.act, a.wprefix, a.alist, a.flags, 0, nullptr,    \
                       a.ctr, a.capA, a.grad_vertices, a.ft, a.pw, n64, n32, a.capF, nullptr,      \
                       a.ubar, a.ubar_e, a.rows, a.contrib);                                       \
    return hipGetLastError();                                                                      \
  }                                                                                                \
  /* the few FP64 cells first: their long per-thread chains start before the FP32 cells' */     \
  /* blocks fill the device */                                                                     \
  hipLaunchKernelGGL((k_bwd_cell<PHASE, double, std::conditional_t<PHASE == 1, double, float>>),   \
                     fgrid, block, 0, s64, a.grids, a.n_grids, a.act_bounds, a.act, a.wprefix,     \
                     a.alist, a.flags, 0, a.fblist, a.ctr, a.capA, a.grad_vertices, a.ft, a.pw,    \
                     n64, n32, a.capF, nullptr, a.ubar, a.ubar_e, a.rows, a.contrib);              \
  {                                                                                                \
    const hipError_t e = hipGetLastError();                                                        \
    if (e != hipSuccess) return e;                                                                 \
  }                                                                                                \
  hipLaunchKernelGGL((k_bwd_cell<PHASE, float, float>), grid, block, 0, s, a.grids, a.n_grids,     \
                     a.act_bounds, a.act, a.wprefix, a.alist, a.flags, 1, nullptr, a.ctr, a.capA,  \
                     a.grad_vertices, a.ft, a.pw, n64, n32, a.capF, a.qstate, a.ubar, a.ubar_e,    \
                     a.rows, a.contrib);                                                           \
  return hipGetLastError();
  if (phase == 1) {
    DC_BWD(1)
  }
  DC_BWD(2)
#undef DC_BWD
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
