// dc_cells.hip — host launchers of the per-face and per-active-cell kernels, exact-arithmetic
// half (built with -ffp-contract=off): pencil words, face normals and the QEF solve of the plain
// FP64 mode (vertices bit-identical to the CPU backend's), quad emission (orientation tests, and
// FP32 mode's near-tie quads oriented again from FP64 vertices solved here) and
// the adjoint's gather. The FP32 mode and the adjoint cell kernels are launched from
// dc_cells_fast.hip.

#include "dc_cells_kernels.h"

// ---------------------------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------------------------
hipError_t dc_launch_face_normals(const DcFaceNormalArgs& a, hipStream_t s) {
  if (a.n_pen <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pencil_words, dim3((unsigned)((a.n_pen + 255) / 256)), dim3(256), 0, s,
                     a.grids, a.n_grids, a.pen_bounds, a.n_pen, a.ft, a.pbase, a.pw);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (!a.list_done) {
    e = dc_launch_face_list(a.ft, a.n_ft, a.tile_sum, a.tile_base, a.flist, a.capF,
                            &a.ctr->n_list, s);
    if (e != hipSuccess) return e;
  }
  // the table is FP64 in both modes (the FP32 mode's FP64 cells and near-tie quads read it like
  // the plain solve); FP32 mode also writes a rounded copy for its FP32 solve and adjoint, whose
  // gathers run faster on half the bytes
  hipLaunchKernelGGL(k_face_normals<double>, dim3((unsigned)a.face_blocks), dim3(256), 0, s,
                     a.grids, a.n_grids, a.ft_bounds, a.flist, a.ft, a.pw, a.ctr, a.capF,
                     reinterpret_cast<double*>(a.nrm), a.nrm32);
  return hipGetLastError();
}

hipError_t dc_launch_qef(const DcQefArgs& a, hipStream_t s) {
  const dim3 grid((unsigned)a.blocks), block(256);
  if (a.precision == DC_PREC_FP64) {
    hipLaunchKernelGGL(k_qef<DC_PREC_FP64>, grid, block, 0, s, a.grids, a.n_grids, a.act_bounds,
                       a.act, a.alist, a.ctr_rw, a.capA, a.v64, a.v32, a.flags, a.fblist, a.ft,
                       a.pw, a.nrm, a.capF, nullptr, nullptr);
    return hipGetLastError();
  }
  return dc_launch_qef_fp32(a, s);
}

hipError_t dc_launch_emit(const DcEmitArgs& a, hipStream_t s) {
  uint64_t* reo = a.precision != DC_PREC_FP64 ? a.reorient : nullptr;
  hipLaunchKernelGGL(k_emit, dim3((unsigned)a.face_blocks), dim3(256), 0, s, a.grids, a.n_grids,
                     a.ft_bounds, a.flist, a.ft, a.pw, a.act, a.wprefix, a.v64, a.v32,
                     a.precision, a.ctr, a.capA, a.capF, a.faces, reo);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !reo) return e;
  // one wave per listed quad (each is four FP64 solves in a row of dependent loads): a quarter of
  // the face blocks, i.e. one wave per 256 quads of capacity (4096 waves at 512^3); waves past the
  // count exit after reading it
  hipLaunchKernelGGL(k_reorient, dim3((unsigned)std::max<int64_t>(1, a.face_blocks / 4)),
                     dim3(256), 0, s, a.grids, a.n_grids, a.ft_bounds, a.ft, a.pw, a.act,
                     a.wprefix, a.ctr, a.capF, reo, a.nrm64, a.faces);
  return hipGetLastError();
}

hipError_t dc_launch_triangulate(const int32_t* faces, uint64_t nf, int n_grids,
                                 const uint32_t* goff, const double* v64, const float* v32,
                                 int use64, int32_t* tris, int64_t blocks, hipStream_t s) {
  if (nf == 0) return hipSuccess;
  hipLaunchKernelGGL(k_triangulate, dim3((unsigned)blocks), dim3(256), 0, s,
                     reinterpret_cast<const int4*>(faces), nf, n_grids, goff, v64, v32, use64,
                     tris);
  return hipGetLastError();
}

hipError_t dc_launch_gather(const DcBwdArgs& a, hipStream_t s) {
  if (a.n_sr <= 0) return hipGetLastError();
  // one wave per row group, or (option "gather_blocks") that many blocks looping over the groups
  const int64_t need = (a.n_sr + 3) / 4;
  const dim3 grid((unsigned)(a.gather_blocks > 0 ? std::min<int64_t>(need, a.gather_blocks) : need));
  const size_t lds = 4 * ((size_t)a.gather_lds_row + 2 * kGatherCap) * 4;
#define DC_GATHER(R)                                                                              \
  hipLaunchKernelGGL(k_bwd_gather<R>, grid, dim3(256), lds, s, a.grids, a.n_grids, a.sr_bounds, \
                     a.n_sr, a.gather_lds_row, a.act, a.wprefix, a.contrib, a.grad_values)
  switch (a.gather_rows) {
    case 1: DC_GATHER(1); break;
    case 2: DC_GATHER(2); break;
    case 4: DC_GATHER(4); break;
    case 8: DC_GATHER(8); break;
    default: return hipErrorInvalidValue;
  }
#undef DC_GATHER
  return hipGetLastError();
}
