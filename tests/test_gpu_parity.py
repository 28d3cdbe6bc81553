"""HIP backend vs the CPU oracle: active set, vertex ids and quads bit-exact; FP64 vertices within
1e-6*h (the tolerance stated for the plain form; the kernels follow the oracle's operation order,
so they are expected to match exactly)."""
import numpy as np
import pytest

import oracle
from grids import small_suite
from contouring_hip import synth

pytestmark = pytest.mark.gpu


def _h(g):
    return min(g.vx, g.vy, g.vz)


def _compare(hip_backend, g, label):
    from contouring_hip import DenseSdfGrid
    print(f"[parity] {label} dims={g.nx}x{g.ny}x{g.nz}", flush=True)
    ref = oracle.run_grid(g)
    grid = DenseSdfGrid(g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz, g.values)
    mesh, stats = hip_backend.extract(grid)
    V64 = mesh.vertices_f64.reshape(-1, 3)
    F = mesh.faces.reshape(-1, 4)
    assert stats.vertex_count == ref.n_vertices, label
    assert stats.active_cells == ref.n_vertices, label
    assert stats.face_count == ref.n_faces, label
    np.testing.assert_array_equal(F, ref.F, err_msg=label)
    if ref.n_vertices:
        err = np.abs(V64 - ref.V64).max()
        assert err <= 1e-6 * _h(g), (label, err)
        np.testing.assert_array_equal(mesh.vertices.reshape(-1, 3), ref.V32, err_msg=label)
    return ref, mesh


@pytest.mark.parametrize("idx", range(len(small_suite())))
def test_parity_small_suite(hip_backend, idx):
    g = small_suite()[idx]
    _compare(hip_backend, g, g.name)


def test_parity_bitexact_fp64_vertices(hip_backend):
    g = synth.csg8(64, seed=1234)
    ref, mesh = _compare(hip_backend, g, "csg8_64")
    np.testing.assert_array_equal(mesh.vertices_f64.reshape(-1, 3), ref.V64)


@pytest.mark.parametrize("n,seed", [(128, 1234), (96, 99)])
def test_parity_csg_medium(hip_backend, n, seed):
    _compare(hip_backend, synth.csg8(n, seed=seed), f"csg8_{n}_s{seed}")


def test_repeat_is_deterministic(hip_backend):
    from contouring_hip import DenseSdfGrid
    g = synth.csg8(80, seed=5)
    grid = DenseSdfGrid(g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz, g.values)
    m1, _ = hip_backend.extract(grid)
    m2, _ = hip_backend.extract(grid)
    np.testing.assert_array_equal(m1.faces, m2.faces)
    np.testing.assert_array_equal(m1.vertices_f64, m2.vertices_f64)


@pytest.mark.parametrize("idx", [0, 1, 3, 6, 10, 13])
def test_triangles_match_oracle(hip_backend, idx):
    """dc_hip_triangulate vs the oracle's triangulate_v1 on the same quads and FP64 vertices."""
    from contouring_hip import DenseSdfGrid
    g = small_suite()[idx]
    print(f"[triangles] {g.name}", flush=True)
    ref = oracle.run_grid(g)
    grid = DenseSdfGrid(g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz, g.values)
    mesh, _ = hip_backend.extract(grid)
    T = hip_backend.triangles().cpu().numpy()
    np.testing.assert_array_equal(T, oracle.triangulate(ref.F, ref.V64), err_msg=g.name)
