"""Error paths and edge cases of the C ABI on the GPU.

Every failure comes back as a status code (raised as DcHipError by the Python layer) and leaves
the context usable; degenerate meshes (no active cell, a single cell) and a grid whose active
cells exceed the first capacity guess (the resize-and-repeat path) still match the oracle.
"""
import ctypes

import numpy as np
import pytest

import oracle
from contouring_hip import DenseSdfGridDevice, HipDualContouringBackend, _lib, synth
from contouring_hip.synth import Grid

pytestmark = pytest.mark.gpu


def _phi(g):
    import torch
    return torch.from_numpy(np.ascontiguousarray(g.values, np.float32)).cuda().reshape(
        g.nz, g.ny, g.nx)


def _extract(be, g, phi=None):
    phi = _phi(g) if phi is None else phi
    return be.extract_tensor(phi, (g.vx, g.vy, g.vz), (g.ox, g.oy, g.oz))


@pytest.mark.parametrize("bad", [float("nan"), float("inf"), float("-inf")])
def test_nonfinite_sample_is_reported_and_context_recovers(bad):
    g = synth.sphere_unit(40)
    vals = np.array(g.values, np.float32)
    vals[40 * 40 * 17 + 40 * 9 + 23] = bad
    bad_g = Grid(vals, 40, g.vx, (g.ox, g.oy, g.oz), "bad")
    be = HipDualContouringBackend("plain-fp64")
    print(f"[errors] non-finite {bad}", flush=True)
    with pytest.raises(_lib.DcHipError) as e:
        _extract(be, bad_g)
    assert e.value.status == _lib.DC_HIP_ERR_NONFINITE
    with pytest.raises(_lib.DcHipError) as e:  # no result after a failed extraction
        be.counts()
    assert e.value.status == _lib.DC_HIP_ERR_NO_FORWARD
    (V, F), _ = _extract(be, g)  # the same context works again
    ref = oracle.run_grid(g)
    assert V.shape[0] == ref.n_vertices and F.shape[0] == ref.n_faces


def test_nonfinite_in_one_grid_of_a_batch():
    g1, g2 = synth.sphere_unit(24), synth.csg8(32, seed=4)
    vals = np.array(g2.values, np.float32)
    vals[100] = np.nan
    dev = [DenseSdfGridDevice(g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz,
                              d_values=_phi(Grid(v, g.nx, g.vx, (g.ox, g.oy, g.oz), "x")).reshape(-1))
           for g, v in ((g1, g1.values), (g2, vals))]
    be = HipDualContouringBackend("fp32-fallback")
    with pytest.raises(_lib.DcHipError) as e:
        be.extract_batch(dev)
    assert e.value.status == _lib.DC_HIP_ERR_NONFINITE


def test_invalid_arguments_and_options():
    lib = _lib.load_hip()
    be = HipDualContouringBackend("plain-fp64")
    ctx = be._ctx
    st = _lib.DcHipStats()
    vals = np.zeros(8, np.float32)
    for dims in ((1, 4, 4), (4, 0, 4), (4, 4, -3), (70000, 2, 2)):
        grid = _lib.DcHipGrid(*dims, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0,
                              vals.ctypes.data_as(ctypes.c_void_p))
        assert lib.dc_hip_extract(ctx, ctypes.byref(grid), 0, ctypes.byref(st)) == \
            _lib.DC_HIP_ERR_INVALID_GRID, dims
    grid = _lib.DcHipGrid(2, 2, 2, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, None)
    assert lib.dc_hip_extract_device(ctx, ctypes.byref(grid), 0, ctypes.byref(st)) == \
        _lib.DC_HIP_ERR_INVALID_ARG
    grid = _lib.DcHipGrid(2, 2, 2, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0,
                          vals.ctypes.data_as(ctypes.c_void_p))
    assert lib.dc_hip_extract(ctx, ctypes.byref(grid), 7, ctypes.byref(st)) == \
        _lib.DC_HIP_ERR_INVALID_ARG  # unknown precision
    for key, value in ((b"mark_rows", 12), (b"mark_layers", 8), (b"mark_waves", 3),
                       (b"gather_rows", 3), (b"gather_blocks", -1), (b"qef_blocks", -1),
                       (b"fallback_state", -2),
                       (b"adjoint_lanes", 2), (b"adjoint_fallback_lanes", 4),
                       (b"no_such_option", 1)):
        assert lib.dc_hip_set_option(ctx, key, value) == _lib.DC_HIP_ERR_INVALID_ARG, key
        assert lib.dc_hip_last_error(ctx)
    # nothing extracted yet: results, triangles and the adjoint report it
    out = ctypes.c_int64()
    assert lib.dc_hip_result_counts(ctx, ctypes.byref(out), ctypes.byref(out)) == \
        _lib.DC_HIP_ERR_NO_FORWARD
    buf = np.zeros(16, np.float32)
    assert lib.dc_hip_backward(ctx, None, buf.ctypes.data_as(ctypes.c_void_p), None) == \
        _lib.DC_HIP_ERR_NO_FORWARD
    assert lib.dc_hip_triangulate(ctx, buf.ctypes.data_as(ctypes.c_void_p)) == \
        _lib.DC_HIP_ERR_NO_FORWARD


def test_empty_mesh_and_zero_gradient():
    import torch
    g = Grid(np.full(20 * 20 * 20, 0.5, np.float32), 20, 0.1, (0.0, 0.0, 0.0), "const")
    be = HipDualContouringBackend("fp32-fallback")
    phi = _phi(g)
    (V, F), st = _extract(be, g, phi)
    assert V.shape == (0, 3) and F.shape == (0, 4) and st.vertex_count == 0
    grad = torch.full_like(phi, float("nan"))
    be.backward(torch.zeros((0, 3), device="cuda"), grad)
    assert torch.count_nonzero(grad).item() == 0
    assert be.triangles().shape == (0, 3)


def test_single_cell_grid():
    vals = np.array([-1, 1, 1, 1, 1, 1, 1, 2], np.float32)  # one negative corner
    g = Grid(vals, 2, 1.0, (0.0, 0.0, 0.0), "cell2")
    be = HipDualContouringBackend("plain-fp64")
    (V, F, V64), _ = be.extract_tensor(_phi(g), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), with_f64=True)
    ref = oracle.run_grid(g)
    assert ref.n_vertices == 1 and V.shape[0] == 1 and F.shape[0] == ref.n_faces == 0
    np.testing.assert_array_equal(V64.cpu().numpy(), ref.V64)


def test_capacity_retry_on_a_checkerboard():
    """Every cell active (47^3 = 103,823 > the first capacity guess of 16,384): the pipeline
    repeats with exact capacities and still matches the oracle."""
    n = 48
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    vals = np.where((i + j + k) % 2 == 0, 1.0, -1.0).astype(np.float32).ravel()
    g = Grid(vals, n, 1.0 / (n - 1), (0.0, 0.0, 0.0), "checker48")
    be = HipDualContouringBackend("plain-fp64")
    print("[errors] checkerboard 48^3", flush=True)
    (V, F, V64), st = be.extract_tensor(_phi(g), (g.vx, g.vy, g.vz), (0.0, 0.0, 0.0),
                                        with_f64=True)
    ref = oracle.run_grid(g)
    assert st.vertex_count == ref.n_vertices == (n - 1) ** 3
    np.testing.assert_array_equal(F.cpu().numpy(), ref.F)
    assert np.abs(V64.cpu().numpy() - ref.V64).max() <= 1e-6 * g.vx
