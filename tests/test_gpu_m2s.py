"""Mesh to SDF on the GPU (dc_hip_mesh_to_sdf) against the oracle's restatement of the reference's
compute_mesh_sdf_device_cuda (src/cuda/mesh_to_sdf_cuda.cu:740-973): every sample bit for bit,
with and without the voxelize option, on closed meshes (icosphere, box, torus, a box with a closed
cavity), a random triangle soup, non-cubic grids with unequal spacing and an offset origin, grids
smaller than one bin, an empty mesh, degenerate triangles, DC meshes fed back (a csg8 extraction
triangulated), device tensors in and out, and the error paths."""
import numpy as np
import pytest

import oracle
from contouring_hip import (HipDualContouringBackend, mesh_to_sdf_device_hip, mesh_to_sdf_hip,
                            synth)
from contouring_hip._lib import (DcHipError, DC_HIP_ERR_INVALID_ARG, DC_HIP_ERR_INVALID_GRID,
                                 DC_HIP_ERR_NONFINITE)

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def be():
    return HipDualContouringBackend("plain-fp64")


def _unit(n):
    h = 1.0 / (n - 1)
    return (n, n, n), (0.0, 0.0, 0.0), (h, h, h)


def _check(be, V, F, dims, origin, spacing, label, voxelize=False, voxel_res=256):
    _say(f"[m2s] {label}: {V.shape[0]} vertices, {F.shape[0]} triangles, grid {dims}, "
         f"voxelize={voxelize} vres={voxel_res}")
    got, st = be.mesh_to_sdf(V, F, *dims, *origin, *spacing, voxelize_first=voxelize,
                             voxel_res=voxel_res)
    want = oracle.mesh_to_sdf(V, F, dims, origin, spacing, voxelize, voxel_res)
    assert got.shape == want.shape
    diff = got.view(np.uint32) != want.view(np.uint32)
    assert not diff.any(), (label, int(diff.sum()), got[diff][:5], want[diff][:5])
    return got, st


MESHES = {
    "icosphere": lambda: synth.icosphere(3),
    "box": lambda: synth.box_mesh(),
    "torus": lambda: synth.torus_mesh(),
    "hollow_box": lambda: synth.hollow_box_mesh(),
    "soup": lambda: synth.random_soup(150),
}


@pytest.mark.parametrize("name", list(MESHES))
@pytest.mark.parametrize("voxelize", [False, True])
def test_meshes_equal_oracle(be, name, voxelize):
    V, F = MESHES[name]()
    _, st = _check(be, V, F, *_unit(44), f"{name} 44^3", voxelize, 64)
    assert st["triangles"] == F.shape[0]
    assert st["active_bins"] > 0


def test_unequal_grid_and_spacing(be):
    V, F = synth.torus_mesh(R=0.3, r=0.08)
    dims, origin, spacing = (37, 50, 29), (-0.05, 0.02, 0.1), (0.031, 0.021, 0.029)
    _check(be, V, F, dims, origin, spacing, "torus 37x50x29")
    _check(be, V, F, dims, origin, spacing, "torus 37x50x29", True, 45)


def test_grid_smaller_than_a_bin(be):
    V, F = synth.icosphere(1, radius=0.3)
    for dims in ((1, 1, 1), (5, 6, 7), (9, 3, 16)):
        h = 1.0 / max(max(dims) - 1, 1)
        _check(be, V, F, dims, (0.1, 0.1, 0.1), (h, h, h), f"icosphere {dims}")
        _check(be, V, F, dims, (0.1, 0.1, 0.1), (h, h, h), f"icosphere {dims}", True, 32)


def test_empty_and_degenerate(be):
    dims, origin, spacing = _unit(20)
    V = np.zeros((0, 3), np.float32)
    F = np.zeros((0, 3), np.int32)
    got, st = _check(be, V, F, dims, origin, spacing, "empty")
    assert (got == 10.0).all() and st["active_bins"] == 0
    _check(be, V, F, dims, origin, spacing, "empty", True, 32)
    # degenerate triangles: a point and a segment (skipped by the evaluation, rasterized by the
    # voxelize option), beside one proper triangle
    V = np.array([[0.5, 0.5, 0.5], [0.5, 0.5, 0.5], [0.5, 0.5, 0.5], [0.2, 0.2, 0.2],
                  [0.4, 0.4, 0.4], [0.6, 0.6, 0.6], [0.3, 0.3, 0.7], [0.7, 0.3, 0.7],
                  [0.3, 0.7, 0.7]], np.float32)
    F = np.arange(9, dtype=np.int32).reshape(3, 3)
    _, st = _check(be, V, F, dims, origin, spacing, "degenerate")
    assert st["triangles"] == 1
    _check(be, V, F, dims, origin, spacing, "degenerate", True, 40)


def test_dc_mesh_fed_back(be):
    """A csg8 extraction, triangulated, sampled again (the CLI's mesh input path)."""
    g = synth.csg8(48, seed=1234)
    from contouring_hip import DenseSdfGrid
    mesh, _ = be.extract(DenseSdfGrid(g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz,
                                      g.values))
    V = mesh.vertices.reshape(-1, 3)
    F = synth.quads_to_triangles(mesh.faces)
    _check(be, V, F, *_unit(48), "csg8 48^3 DC mesh")
    _check(be, V, F, *_unit(48), "csg8 48^3 DC mesh", True, 80)


def test_device_tensors_and_module_functions(be):
    import torch
    V, F = synth.icosphere(2)
    dims, origin, spacing = _unit(33)
    want = oracle.mesh_to_sdf(V, F, dims, origin, spacing)
    Vt = torch.from_numpy(V).cuda()
    Ft = torch.from_numpy(F).cuda()
    got, _ = be.mesh_to_sdf(Vt, Ft, *dims, *origin, *spacing, out="device")
    assert got.is_cuda and tuple(got.shape) == (33, 33, 33)
    assert np.array_equal(got.cpu().numpy().view(np.uint32), want.view(np.uint32))
    out = torch.full((33 * 33 * 33,), 7.0, device="cuda")
    be.mesh_to_sdf(V, F, *dims, *origin, *spacing, out=out)
    assert np.array_equal(out.cpu().numpy().reshape(want.shape).view(np.uint32),
                          want.view(np.uint32))
    grid, ms = mesh_to_sdf_hip(V, F, *dims, *origin, *spacing)
    assert ms > 0 and grid.nx == 33
    assert np.array_equal(np.asarray(grid.values).reshape(want.shape).view(np.uint32),
                          want.view(np.uint32))
    dgrid, _ = mesh_to_sdf_device_hip(V, F, *dims, *origin, *spacing)
    assert np.array_equal(dgrid.d_values.cpu().numpy().reshape(want.shape).view(np.uint32),
                          want.view(np.uint32))
    # the device grid goes straight into the extraction
    mesh, st = be.extract_device(dgrid)
    assert st.vertex_count > 0


def test_errors(be):
    V, F = synth.icosphere(1)
    dims, origin, spacing = _unit(16)
    bad = F.copy()
    bad[3, 1] = V.shape[0]
    with pytest.raises(DcHipError) as e:
        be.mesh_to_sdf(V, bad, *dims, *origin, *spacing)
    assert e.value.status == DC_HIP_ERR_INVALID_ARG
    bad[3, 1] = -1
    with pytest.raises(DcHipError) as e:
        be.mesh_to_sdf(V, bad, *dims, *origin, *spacing)
    assert e.value.status == DC_HIP_ERR_INVALID_ARG
    Vn = V.copy()
    Vn[5, 2] = np.nan
    with pytest.raises(DcHipError) as e:
        be.mesh_to_sdf(Vn, F, *dims, *origin, *spacing)
    assert e.value.status == DC_HIP_ERR_NONFINITE
    with pytest.raises(DcHipError) as e:
        be.mesh_to_sdf(V, F, 0, 16, 16, *origin, *spacing)
    assert e.value.status == DC_HIP_ERR_INVALID_GRID
    with pytest.raises(DcHipError) as e:
        be.mesh_to_sdf(V, F, *dims, *origin, 0.1, -0.1, 0.1)
    assert e.value.status == DC_HIP_ERR_INVALID_GRID
    # the context still works afterwards
    _check(be, V, F, dims, origin, spacing, "after errors")
