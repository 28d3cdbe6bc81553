"""MI355X dual contouring of dense SDF grids — drop-in for the reference's dense GPU backend and its
sparse brick / multi-vertex backends.

Mirrors the reference Python module (src/python/contouring, src/cpp/bindings.cpp): grid, mesh
and stats classes plus a backend class whose extract(grid) returns (mesh, stats). It is a ctypes
package over libdc_hip.so with its own name, so it installs beside the reference's `contouring`
package; inside the reference's own module the same class is bound with pybind11 by
include/dc_hip_pybind.h.
"""
from .backend import (  # noqa: F401
    DenseSdfGrid,
    DenseSdfGridDevice,
    DualContouringMesh,
    DualContouringStats,
    HipDualContouringBackend,
    HipSparseDualContouringBackend,
    HipSparseMvdcDualContouringBackend,
    NormalComputationMode,
    PRECISIONS,
    mesh_to_sdf_device_hip,
    mesh_to_sdf_hip,
    postprocess_mesh,
)
from . import synth  # noqa: F401
from .io import load_obj, load_sdf, save_obj, save_sdf  # noqa: F401
from .autotune import autotune_launch, autotune_mark  # noqa: F401
