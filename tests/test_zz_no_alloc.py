"""dc_hip_extract_device_noalloc (every intermediate in the library's static device scratch, the
mesh in the caller's buffers) against dc_hip_extract_device on the same grids, called through the
C ABI: vertex and face counts, quads and FP32 vertices bit-identical (FP64 vertices too in plain
mode), on two calls in a row whose second grid is smaller; an output capacity one short returns
DC_HIP_ERR_INVALID_ARG with the counts and writes nothing past the capacities; a grid beyond the
scratch is rejected before anything runs.
"""
import ctypes

import numpy as np
import pytest

from contouring_hip import DenseSdfGridDevice, HipDualContouringBackend, _lib, synth
from contouring_hip.backend import PRECISIONS

pytestmark = pytest.mark.gpu


def _dev(g):
    import torch
    phi = torch.from_numpy(np.ascontiguousarray(g.values, np.float32)).cuda()
    return phi, DenseSdfGridDevice(g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz,
                                   d_values=phi)


def _noalloc(be, g, phi, vcap, fcap, with64, pad=8):
    """Runs the entry point into fresh buffers with `pad` sentinel rows past each capacity."""
    import torch
    be._use_torch_stream()
    V = torch.full((vcap + pad, 3), -7.0, device="cuda")
    V64 = torch.full((vcap + pad, 3), -7.0, dtype=torch.float64, device="cuda") if with64 else None
    F = torch.full((fcap + pad, 4), -7, dtype=torch.int32, device="cuda")
    nv, nf = ctypes.c_int64(-1), ctypes.c_int64(-1)
    cg = be._cgrid(g, phi.data_ptr())
    rc = be._lib.dc_hip_extract_device_noalloc(
        be._ctx, ctypes.byref(cg), PRECISIONS[be.precision], ctypes.c_void_p(V.data_ptr()),
        ctypes.c_void_p(V64.data_ptr() if with64 else 0), vcap, ctypes.c_void_p(F.data_ptr()),
        fcap, ctypes.byref(nv), ctypes.byref(nf))
    torch.cuda.synchronize()
    return rc, nv.value, nf.value, V, V64, F


@pytest.mark.parametrize("precision", ["plain-fp64", "fp32-fallback"])
def test_noalloc_matches_extract_device_twice(precision):
    import torch
    be = HipDualContouringBackend(precision)
    ref = HipDualContouringBackend(precision)
    plain = precision == "plain-fp64"
    for n, seed in ((96, 1234), (48, 7)):  # the second call's grid is smaller
        g = synth.csg8(n, seed=seed)
        print(f"[noalloc] {precision} csg8 {n}", flush=True)
        phi, dg = _dev(g)
        (V_ref, F_ref, *rest), _ = ref.extract_tensor(
            phi.reshape(g.nz, g.ny, g.nx), (g.vx, g.vy, g.vz), (g.ox, g.oy, g.oz),
            with_f64=plain)
        nv_ref, nf_ref = V_ref.shape[0], F_ref.shape[0]
        rc, nv, nf, V, V64, F = _noalloc(be, g, phi, nv_ref + 100, nf_ref + 100, plain)
        assert rc == _lib.DC_HIP_OK, be._lib.dc_hip_last_error(be._ctx)
        assert (nv, nf) == (nv_ref, nf_ref)
        assert torch.equal(V[:nv], V_ref)
        assert torch.equal(F[:nf], F_ref)
        if plain:
            assert torch.equal(V64[:nv], rest[0])
        assert bool((V[nv:] == -7.0).all()) and bool((F[nf:] == -7).all())


def test_noalloc_capacity_too_small_reports_counts():
    be = HipDualContouringBackend("plain-fp64")
    g = synth.csg8(40, seed=3)
    phi, dg = _dev(g)
    print("[noalloc] capacities", flush=True)
    rc, nv, nf, V, V64, F = _noalloc(be, g, phi, 1 << 16, 1 << 17, True)
    assert rc == _lib.DC_HIP_OK and nv > 1 and nf > 1
    for vcap, fcap in ((nv - 1, nf), (nv, nf - 1)):
        rc2, nv2, nf2, V2, V642, F2 = _noalloc(be, g, phi, vcap, fcap, True)
        assert rc2 == _lib.DC_HIP_ERR_INVALID_ARG
        assert (nv2, nf2) == (nv, nf)
        assert bool((V2[vcap:] == -7.0).all()) and bool((F2[fcap:] == -7).all())
        assert bool((V642[vcap:] == -7.0).all())
    # exact capacities succeed
    rc3, nv3, nf3, V3, _, F3 = _noalloc(be, g, phi, nv, nf, True)
    assert rc3 == _lib.DC_HIP_OK and (nv3, nf3) == (nv, nf)


def test_noalloc_rejects_grid_beyond_scratch():
    import torch
    be = HipDualContouringBackend("plain-fp64")
    n = 512
    phi = torch.ones(n * n * n, device="cuda")  # a real grid: nothing is read before the check
    g = DenseSdfGridDevice(n, n, n, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, d_values=phi)
    print("[noalloc] grid beyond the scratch", flush=True)
    rc, nv, nf, *_ = _noalloc(be, g, phi, 16, 16, False)
    assert rc == _lib.DC_HIP_ERR_INVALID_GRID
    assert b"static scratch" in be._lib.dc_hip_last_error(be._ctx)
    # the context still works afterwards, and so does the entry point on a small grid
    small = synth.csg8(24, seed=9)
    sphi, _ = _dev(small)
    rc, nv, nf, *_ = _noalloc(be, small, sphi, 1 << 14, 1 << 15, False)
    assert rc == _lib.DC_HIP_OK and nv > 0
