"""Sparse brick dual contouring and its multi-vertex form on the GPU (dc_hip_extract_sparse[_device])
against the oracle's restatement of the reference's CudaSparseDualContouringBackend /
CudaSparseMvdcDualContouringBackend (src/cuda/dc_cuda_sparse.cu:1170-1485): every vertex bit for
bit, every quad, and every counter of the stats, over the dense suite's grids and grids made for
this path (thin sheets, near-zero noise that fills the 12-entry cap, crossed sheets, exact-zero
regions whose QEF falls back to the mass point, a tilted plane, a non-cubic grid with unequal
spacing), brick sizes 1 to beyond the grid, the three normal modes, device input and device
output, and the error paths."""
import numpy as np
import pytest

import grids
import oracle
from contouring_hip import (DenseSdfGrid, DenseSdfGridDevice, HipSparseDualContouringBackend,
                            HipSparseMvdcDualContouringBackend, NormalComputationMode, synth)
from contouring_hip._lib import (DcHipError, DC_HIP_ERR_INVALID_ARG, DC_HIP_ERR_INVALID_GRID,
                                 DC_HIP_ERR_NONFINITE)

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def backends():
    return {False: HipSparseDualContouringBackend(), True: HipSparseMvdcDualContouringBackend()}


def _dense(g):
    return DenseSdfGrid(g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz, g.values)


def _check(backends, g, mv, brick=8, mode=0):
    _say(f"[sparse] {g.name} {g.nx}x{g.ny}x{g.nz} mv={mv} brick={brick} mode={mode}")
    be = backends[mv]
    be.brick_size, be.normal_mode = brick, mode
    mesh, st = be.extract(_dense(g))
    want = oracle.sparse_grid(g, brick_size=brick, normal_mode=mode, multi_vertex=mv)
    V = np.asarray(mesh.vertices, np.float32).reshape(-1, 3)
    F = np.asarray(mesh.faces, np.int32).reshape(-1, 4)
    assert V.shape == want["vertices"].shape, (g.name, V.shape, want["vertices"].shape)
    diff = V.view(np.uint32) != want["vertices"].view(np.uint32)
    assert not diff.any(), (g.name, int(diff.any(axis=1).sum()), np.argwhere(diff)[:4])
    assert F.shape == want["faces"].shape, (g.name, F.shape, want["faces"].shape)
    assert np.array_equal(F, want["faces"]), (g.name, np.argwhere(F != want["faces"])[:4])
    for k, v in want["stats"].items():
        assert getattr(st, k) == v, (g.name, k, getattr(st, k), v)
    assert st.backend == ("hip-sparse-mvdc" if mv else "hip-sparse")
    return mesh, st, want


SUITE = grids.sparse_suite()


@pytest.mark.parametrize("g", SUITE, ids=[g.name for g in SUITE])
@pytest.mark.parametrize("mv", [False, True], ids=["sparse", "mvdc"])
def test_suite_equals_oracle(backends, g, mv):
    _check(backends, g, mv)


@pytest.mark.parametrize("brick", [1, 3, 5, 16, 64, 1000])
def test_brick_sizes(backends, brick):
    for g in (grids.slab_grid(24), grids.crossed_sheets(28), synth.csg8(40, seed=3),
              grids.near_noise_grid(13, 9, neg=0.4)):
        for mv in (False, True):
            _check(backends, g, mv, brick=brick)


@pytest.mark.parametrize("mode", [NormalComputationMode.FiniteDifference,
                                  NormalComputationMode.EdgeGradient,
                                  NormalComputationMode.PrecomputedGradient])
def test_normal_modes(backends, mode):
    for g in (synth.csg8(48, seed=21), grids.near_noise_grid(20, 5), grids.uneven_blob(),
              grids.crossed_sheets(28)):
        for mv in (False, True):
            _check(backends, g, mv, brick=8 if g.nx > 30 else 100, mode=mode)


def test_thin_sheet_splits(backends):
    """One brick over the whole slab grid: every cell of the sheet layer holds four thin z edges
    and splits into two vertices (checked against the oracle; the geometry is pinned on the CPU)."""
    _, st, _ = _check(backends, grids.slab_grid(24), True, brick=100)
    assert st.two_vertex_cells > 0 and st.vertex_count > st.active_cells


def test_larger_grids(backends):
    for g in (synth.csg8(128, seed=1234), synth.blend64(96, seed=9)):
        for mv in (False, True):
            _check(backends, g, mv)


def test_device_input_and_output(backends):
    import torch
    g = synth.csg8(64, seed=77)
    want = oracle.sparse_grid(g, multi_vertex=True)
    t = torch.from_numpy(np.asarray(g.values, np.float32).copy()).cuda()
    dg = DenseSdfGridDevice(g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz, t)
    be = backends[True]
    be.brick_size, be.normal_mode = 8, 0
    mesh, st = be.extract_device(dg)
    assert np.array_equal(np.asarray(mesh.vertices).view(np.uint32),
                          want["vertices"].ravel().view(np.uint32))
    assert np.array_equal(np.asarray(mesh.faces), want["faces"].ravel())
    (V, F), st2 = be.extract_device(dg, download=False)
    assert V.is_cuda and F.is_cuda
    assert np.array_equal(V.cpu().numpy().view(np.uint32), want["vertices"].view(np.uint32))
    assert np.array_equal(F.cpu().numpy(), want["faces"])
    assert st2.vertex_count == want["stats"]["vertex_count"]
    # an int device pointer works too
    dg.d_values = t.data_ptr()
    mesh, _ = be.extract_device(dg)
    assert np.array_equal(np.asarray(mesh.faces), want["faces"].ravel())


def test_sizes_and_empty(backends):
    # smallest grid, all-positive grid (no active brick), ragged sizes across 64-bit words
    for g in (grids.ref_sphere(2), grids.ref_const(16, 1.0), grids.ragged_sphere(130, 9, 70),
              grids.ragged_sphere(65, 66, 3), grids.ragged_sphere(3, 129, 5)):
        for mv in (False, True):
            _, st, _ = _check(backends, g, mv)
    mesh, st = backends[False].extract(_dense(grids.ref_const(16, -1.0)))
    assert st.active_bricks == 0 and st.vertex_count == 0 and len(mesh.faces) == 0


def test_errors(backends):
    be = backends[False]
    g = synth.sphere_unit(16)
    bad = np.asarray(g.values, np.float32).copy()
    bad[100] = np.nan
    with pytest.raises(DcHipError) as e:
        be.extract(DenseSdfGrid(16, 16, 16, g.vx, g.vx, g.vx, 0, 0, 0, bad))
    assert e.value.status == DC_HIP_ERR_NONFINITE
    bad[100] = np.inf
    with pytest.raises(DcHipError) as e:
        be.extract(DenseSdfGrid(16, 16, 16, g.vx, g.vx, g.vx, 0, 0, 0, bad))
    assert e.value.status == DC_HIP_ERR_NONFINITE
    with pytest.raises(DcHipError) as e:
        be.extract(DenseSdfGrid(1, 16, 256, g.vx, g.vx, g.vx, 0, 0, 0, g.values))
    assert e.value.status == DC_HIP_ERR_INVALID_GRID
    for brick, mode in ((0, 0), (-3, 0), (8, 3), (8, -1)):
        be.brick_size, be.normal_mode = brick, mode
        with pytest.raises(DcHipError) as e:
            be.extract(_dense(g))
        assert e.value.status == DC_HIP_ERR_INVALID_ARG
    be.brick_size, be.normal_mode = 8, 0
    # the context still works afterwards
    _check(backends, g, False)


@pytest.mark.timeout(240)
def test_large_host_grid_matches_device_input(backends):
    """A host grid of 128 MiB or more goes up through the pinned, pipelined upload (then the device
    path): the same mesh and counters as the device-resident grid, bit for bit."""
    import torch
    g = synth.csg8(384, seed=1234)  # 226 MB of samples: four 64 MiB chunks
    t = torch.from_numpy(np.asarray(g.values, np.float32).copy()).cuda()
    dg = DenseSdfGridDevice(g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz, t)
    for mv in (False, True):
        _say(f"[sparse] host vs device input, csg8 384^3 mv={mv}")
        be = backends[mv]
        be.brick_size, be.normal_mode = 8, 0
        m_host, st_host = be.extract(_dense(g))
        m_dev, st_dev = be.extract_device(dg)
        assert np.array_equal(np.asarray(m_host.vertices).view(np.uint32),
                              np.asarray(m_dev.vertices).view(np.uint32))
        assert np.array_equal(np.asarray(m_host.faces), np.asarray(m_dev.faces))
        assert st_host.vertex_count == st_dev.vertex_count > 0
        assert st_host.face_count == st_dev.face_count
        assert st_host.upload_ms > 0
