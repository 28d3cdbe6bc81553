// The reference dc_cli's backend factory (src/cpp/dc_cli.cpp:310-329) with the branches the
// drop-in adds (dense, sparse, multi-vertex), and the device-grid call site; compiled (syntax
// only) against the reference's own headers, with and without DC_ENABLE_CUDA, by
// tests/test_drop_in.py.
#include <memory>
#include <string>

#include "dc_backend.h"
#ifdef DC_ENABLE_CUDA
#include "dc_cuda.h"
#include "mesh_to_sdf.h"
#endif
#include "dc_hip_backend.h"

std::unique_ptr<IDualContouringBackend> make_backend(const std::string& backend_type) {
  std::unique_ptr<IDualContouringBackend> backend;
  if (backend_type == "hip") {
    backend = std::make_unique<HipDualContouringBackend>(/*device=*/0, DC_HIP_PLAIN_FP64);
  } else if (backend_type == "hip-sparse") {
    backend = std::make_unique<HipSparseDualContouringBackend>(8, DC_HIP_NORMAL_FINITE_DIFFERENCE);
  } else if (backend_type == "hip-sparse-mvdc" || backend_type == "mvdc") {
    backend = std::make_unique<HipSparseMvdcDualContouringBackend>(8, DC_HIP_NORMAL_EDGE_GRADIENT);
#ifdef DC_ENABLE_CUDA
  } else if (backend_type == "cuda" || backend_type == "cuda-dense") {
    backend = std::make_unique<CudaDualContouringBackend>();
#endif
  } else {
    backend = std::make_unique<CpuDualContouringBackend>();
  }
  return backend;
}

DualContouringMesh run(IDualContouringBackend& b, const DenseSdfGrid& grid) {
  DualContouringStats stats;
  DualContouringMesh mesh = b.extract(grid, stats);
  return postprocess_mesh(mesh, /*close_holes=*/true, /*remove_floaters=*/true);
}

DualContouringMesh run_device(const DenseSdfGridDevice& g) {
  HipDualContouringBackend hip(0, DC_HIP_FP32_FALLBACK);
  DualContouringStats st;
  return hip.extract_device(g, st);
}
