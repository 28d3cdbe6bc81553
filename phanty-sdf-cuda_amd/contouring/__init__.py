"""MI355X dual contouring of dense SDF grids — drop-in for the reference's dense GPU backend.

Mirrors the reference Python module (src/python/contouring, src/cpp/bindings.cpp): grid, mesh
and stats classes plus a backend class whose extract(grid) returns (mesh, stats).
"""
from .backend import (  # noqa: F401
    DenseSdfGrid,
    DenseSdfGridDevice,
    DualContouringMesh,
    DualContouringStats,
    HipDualContouringBackend,
    PRECISIONS,
)
from . import synth  # noqa: F401
from .autotune import autotune_mark  # noqa: F401
