"""Pins for the CPU oracle (the restatement of the reference CpuDualContouringBackend).

The reference ships no golden vectors (its dc_tests.cpp checks counts only), so the oracle is
pinned by: an independently written NumPy twin that follows the reference's data structures and
uses a library SVD; closed forms (planes, boxes, empty/full grids); exhaustive single-cell sign
patterns; library eigen-solvers for the 3x3 solve; Euler characteristic of closed meshes; and
committed fixtures that freeze the verified outputs.
"""
import itertools
import os

import numpy as np
import pytest

import oracle
from oracle import twin
from grids import ref_box, ref_const, ref_plane, ref_sphere, ref_thin_gap, ragged_sphere, \
    zeros_grid
from contouring_hip import synth

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _twin(g):
    return twin.extract(g.values, g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz)


TWIN_GRIDS = [
    lambda: ref_sphere(12), lambda: ref_box(11), lambda: ref_plane(7), lambda: ref_thin_gap(13),
    lambda: zeros_grid(), lambda: zeros_grid(5, 6, 4, seed=9), lambda: ragged_sphere(10, 7, 9),
    lambda: synth.csg8(12, seed=3), lambda: synth.sphere_unit(11),
]


@pytest.mark.parametrize("make", TWIN_GRIDS)
def test_oracle_matches_independent_twin(make):
    g = make()
    r = oracle.run_grid(g)
    V, F, act = _twin(g)
    np.testing.assert_array_equal(r.active, act)
    np.testing.assert_array_equal(r.F, F)
    if len(V):
        # library SVD vs the oracle's Jacobi eigen-solver: rounding-level differences only
        assert np.abs(r.V64 - V).max() <= 1e-9 * g.vx


def test_empty_and_full_grids_have_no_mesh():
    for v in (1.0, -1.0, 0.0):
        r = oracle.run_grid(ref_const(16, v))
        assert r.n_vertices == 0 and r.n_faces == 0


def test_plane_closed_form():
    # value z on [-1,1]^3, 16^3: the zero level z = 0 lies strictly inside one cell layer
    g = ref_plane(16)
    r = oracle.run_grid(g)
    assert r.n_vertices == 15 * 15
    k = r.active // (15 * 15)
    assert np.all(k == k[0])
    # rank-1 QEF: the vertex lies on the plane, x and y at the centroid of the hermite points
    assert np.abs(r.V64[:, 2]).max() < 1e-12
    # interior z-edges crossing: (cx-1)*(cy-1) quads
    assert r.n_faces == 14 * 14


def test_single_cell_sign_patterns_exhaustive():
    """has_vertex (Cell.cpp:66-76) over all 3^8 sign/zero patterns of a 2x2x2 grid."""
    for pat in itertools.product((-1.0, 0.0, 1.0), repeat=8):
        vals = np.array(pat, np.float32)
        r = oracle.run(vals, 2, 2, 2, 1.0, 1.0, 1.0)
        expect = int(vals.min() < 0 and vals.max() > 0)
        assert r.n_vertices == expect, pat
        assert r.n_faces == 0  # no interior edges in a single cell


def test_sym3_eig_matches_numpy():
    rng = np.random.default_rng(0)
    for _ in range(300):
        A = rng.normal(size=(rng.integers(1, 13), 3))
        A /= np.linalg.norm(A, axis=1, keepdims=True)
        M = A.T @ A
        lam, Q = oracle.sym3_eig(M)
        np.testing.assert_allclose(np.sort(lam), np.linalg.eigvalsh(M), atol=1e-12)
        np.testing.assert_allclose(Q @ np.diag(lam) @ Q.T, M, atol=1e-12)
        np.testing.assert_allclose(Q.T @ Q, np.eye(3), atol=1e-12)


def test_qef_solve_matches_truncated_pinv():
    """Cell.cpp:324-361: x = c + V diag(1/s, s > 0.01) U^T (A^T b - A^T A c)."""
    rng = np.random.default_rng(1)
    for _ in range(300):
        m = int(rng.integers(1, 13))
        n = rng.normal(size=(m, 3))
        if rng.uniform() < 0.3:  # near rank-deficient clusters
            n[:, 2] *= 1e-3
        n /= np.linalg.norm(n, axis=1, keepdims=True)
        p = rng.uniform(-1, 1, size=(m, 3))
        x = oracle.qef_solve(n, p)
        b = np.einsum("ij,ij->i", n, p)
        M, r = n.T @ n, n.T @ b
        U, S, Vt = np.linalg.svd(M)
        if np.any(np.abs(S - 0.01) < 1e-9):
            continue
        P = Vt.T @ np.diag([1 / s if s > 0.01 else 0.0 for s in S]) @ U.T
        c = p.mean(axis=0)
        np.testing.assert_allclose(x, c + P @ (r - M @ c), atol=1e-9)


def test_box_face_vertices_lie_on_the_faces():
    """Axis-aligned box |x|=0.5: away from its edges the SDF is linear across a face, the
    hermite normals are the face normal and the rank-1 QEF puts the vertex on the face."""
    g = ref_box(32)
    r = oracle.run_grid(g)
    h = g.vx
    on_face = 0
    for v in r.V64:
        a = np.abs(v)
        d = int(np.argmax(a))
        others = np.delete(a, d)
        if others.max() < 0.5 - 2 * h:
            assert abs(a[d] - 0.5) < 1e-9, v
            on_face += 1
    assert on_face > 100


def test_closed_sphere_mesh_is_manifold_with_euler_2():
    g = ref_sphere(32)
    r = oracle.run_grid(g)
    F = r.F
    edges = {}
    for q in F:
        for a, b in zip(q, np.roll(q, -1)):
            key = (min(a, b), max(a, b))
            edges[key] = edges.get(key, 0) + 1
    assert all(v == 2 for v in edges.values())
    assert r.n_vertices - len(edges) + len(F) == 2


def test_orientation_agrees_with_gradient():
    g = synth.sphere_unit(24)
    r = oracle.run_grid(g)
    V = r.V64
    for q in r.F:
        pa, pb, pd = V[q[0]], V[q[1]], V[q[3]]
        n = np.cross(pb - pa, pd - pa)
        c = V[q].mean(axis=0) - 0.5
        assert np.dot(n, c) > 0  # outward for a sphere whose SDF grows outward


def test_hermite_normals_averaged_consistently():
    """After averaging, the 4 cells around an averaged edge hold the same normal."""
    g = synth.csg8(14, seed=2)
    r = oracle.run_grid(g)
    mask, pts, nrm = r.hermite(averaged=True)
    cx = cy = g.nx - 1
    vid = {int(c): v for v, c in enumerate(r.active)}
    checked = 0
    for v, c in enumerate(r.active):
        i, j, k = c % cx, (c // cx) % cy, c // (cx * cy)
        # X edge 0 of this cell is edge 2 of (i,j-1,k), 4 of (i,j,k-1), 6 of (i,j-1,k-1)
        if not (mask[v] & 1) or j < 1 or k < 1:
            continue
        group = [(i, j, k, 0), (i, j - 1, k, 2), (i, j, k - 1, 4), (i, j - 1, k - 1, 6)]
        ids = [vid.get(a + cx * (b + cy * cc)) for a, b, cc, _ in group]
        if None in ids:
            continue
        ns = [nrm[w, e] for w, (_, _, _, e) in zip(ids, group)]
        for n in ns[1:]:
            np.testing.assert_array_equal(n, ns[0])
        checked += 1
    assert checked > 10


def test_fixture_regression():
    """Committed outputs of the verified oracle (tests/golden/make_fixtures.py)."""
    path = os.path.join(GOLDEN, "oracle_fixtures.npz")
    data = np.load(path, allow_pickle=False)
    names = sorted({k.split("__")[0] for k in data.files})
    assert names
    for name in names:
        vals = data[f"{name}__values"]
        dims = data[f"{name}__dims"]
        geo = data[f"{name}__geo"]
        r = oracle.run(vals, *[int(d) for d in dims], *[float(x) for x in geo])
        np.testing.assert_array_equal(r.F, data[f"{name}__faces"], err_msg=name)
        np.testing.assert_array_equal(r.active, data[f"{name}__active"], err_msg=name)
        np.testing.assert_allclose(r.V64, data[f"{name}__v64"], atol=1e-12, err_msg=name)
