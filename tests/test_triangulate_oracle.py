"""The oracle's quad -> triangle rule (triangulate_v1, contouring.cpp:88-113, with
orient_triangles_for_quad, contouring.cpp:40-86), pinned by closed forms and mesh invariants."""
import numpy as np

import oracle
from grids import ref_sphere

Q = np.array([[0, 1, 2, 3]], np.int32)


def test_convex_square_keeps_both_triangles():
    V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], float)
    np.testing.assert_array_equal(oracle.triangulate(Q, V), [[0, 1, 2], [0, 2, 3]])
    # the reverse winding is its own quad normal: still no flip
    np.testing.assert_array_equal(oracle.triangulate(Q, V[[0, 3, 2, 1]]), [[0, 1, 2], [0, 2, 3]])


def test_folded_quad_flips_both_triangles():
    # c behind a: both triangles' normals point against (b-a) x (d-a) = +z
    V = np.array([[0, 0, 0], [2, 0, 0], [-1, -1, 0], [0, 2, 0]], float)
    np.testing.assert_array_equal(oracle.triangulate(Q, V), [[0, 2, 1], [0, 3, 2]])


def test_collinear_abd_uses_the_other_diagonal():
    # (b-a) x (d-a) = 0, so the quad normal is (c-a) x (d-a) = -z: triangle (a,b,c) (+z) flips
    V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [2, 0, 0]], float)
    np.testing.assert_array_equal(oracle.triangulate(Q, V), [[0, 2, 1], [0, 2, 3]])


def test_degenerate_quad_is_left_as_split():
    V = np.zeros((4, 3))
    np.testing.assert_array_equal(oracle.triangulate(Q, V), [[0, 1, 2], [0, 2, 3]])


def test_sphere_mesh_invariants():
    g = ref_sphere(24)
    r = oracle.run_grid(g)
    T = oracle.triangulate(r.F, r.V64)
    assert T.shape == (2 * r.n_faces, 3)
    V = r.V64
    for f, (a, b, c, d) in enumerate(r.F):
        t0, t1 = T[2 * f], T[2 * f + 1]
        # the pair covers the quad, sharing the diagonal (a, c), and keeps a as first vertex
        assert t0[0] == a and t1[0] == a
        assert sorted(t0) == sorted([a, b, c]) and sorted(t1) == sorted([a, c, d])
        qn = np.cross(V[b] - V[a], V[d] - V[a])
        if qn @ qn < 1e-24:
            qn = np.cross(V[c] - V[a], V[d] - V[a])
        if qn @ qn < 1e-24:
            continue
        for t in (t0, t1):
            tn = np.cross(V[t[1]] - V[t[0]], V[t[2]] - V[t[0]])
            assert tn @ qn >= 0.0
