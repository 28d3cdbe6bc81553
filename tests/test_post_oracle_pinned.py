"""The post-processing oracle (oracle/post_oracle.cpp, the reference's postprocess_mesh,
src/cpp/dc_backend.cpp:127-558) pinned by hand-derived answers, brute force and invariants (CPU).

- weld: the welded vertices are the first vertex of each distinct key round(x / eps) (half away
  from zero) with eps = 1e-5 x the box diagonal, all in FP32, computed here with NumPy; faces with
  fewer than 3 distinct ids dropped;
- floaters: the kept faces are the largest connected component (SciPy) of the graph joining faces
  that share an edge used by exactly two face sides; ties keep the component of the lowest face;
  an edge used by three faces joins none of them;
- holes: hand-walked loops of small meshes (a quad, two quads sharing a vertex); a frustum of
  regular octagons, whose fill steps each ring a mean edge length towards the centroid (closed-form
  radii); loops longer than 4096 left open; a sphere cut open by the grid becomes a closed,
  consistently oriented surface of Euler characteristic 2;
- output: vertices in order of first use (NumPy's first-occurrence order).
"""
import numpy as np
import pytest

import oracle


def _half_away(x):
    return np.sign(x) * np.floor(np.abs(x) + np.float32(0.5))


def _edges(F):
    """Directed sides with distinct ends of every face."""
    out = []
    for f in F.tolist():
        for e in range(4):
            u, v = f[e], f[(e + 1) % 4]
            if u != v:
                out.append((u, v))
    return out


def _first_use(F):
    """Renumbers F's ids by first use: (new F, old ids in new order)."""
    flat = F.ravel()
    _, first = np.unique(flat, return_index=True)
    order = flat[np.sort(first)]
    new = np.empty(order.max() + 1, np.int64)
    new[order] = np.arange(order.size)
    return new[F].astype(np.int32), order


def test_single_quad_becomes_a_closed_pillow():
    # the quad's boundary 0->1->2->3->0 is one loop; walked from vertex 0 it is [0,1,2,3] and is
    # filled reversed: one quad (3,2,1,0)
    V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    F = np.array([[0, 1, 2, 3]], np.int32)
    V2, F2, info = oracle.postprocess(V, F, True, True)
    np.testing.assert_array_equal(V2, V)
    np.testing.assert_array_equal(F2, [[0, 1, 2, 3], [3, 2, 1, 0]])
    assert info["loops_filled"] == 1 and info["boundary_edges"] == 4
    # neither flag: the input as it is; floaters only: the same single quad
    V3, F3, _ = oracle.postprocess(V, F, False, False)
    np.testing.assert_array_equal(F3, F)
    V4, F4, _ = oracle.postprocess(V, F, False, True)
    np.testing.assert_array_equal(F4, F)


def test_two_quads_sharing_a_vertex_cut_a_loop_off_the_walk():
    # Boundary lists (sides sorted by (min, max)): 0:[1] 1:[2] 2:[3,4] 3:[0] 4:[5] 5:[6] 6:[2].
    # The walk from 0 goes 0,1,2, takes 2's last edge to 4, then 5, 6 and meets 2 again: the loop
    # [2,4,5,6] is cut off and filled reversed (6,5,4,2); the walk resumes at 2 -> 3 -> 0: closed,
    # [0,1,2,3] filled reversed (3,2,1,0).
    V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [2, 1, 0], [2, 2, 0], [1, 2, 0]],
                 np.float32)
    F = np.array([[0, 1, 2, 3], [2, 4, 5, 6]], np.int32)
    V2, F2, info = oracle.postprocess(V, F, True, False)
    np.testing.assert_array_equal(V2, V)
    np.testing.assert_array_equal(F2, [[0, 1, 2, 3], [2, 4, 5, 6], [6, 5, 4, 2], [3, 2, 1, 0]])
    assert info["loops_filled"] == 2 and info["added_vertices"] == 0


def test_edge_of_three_faces_joins_none_and_ties_keep_the_lowest_face():
    # a book of three quads on edge (0,1): with floater removal that edge links no faces, so each
    # quad is its own component of size 1; the first (face 0) is kept
    V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [1, -1, 0], [0, -1, 0], [1, 0, 1],
                  [0, 0, 1]], np.float32)
    F = np.array([[0, 1, 2, 3], [1, 0, 5, 4], [0, 1, 6, 7]], np.int32)
    V2, F2, info = oracle.postprocess(V, F, False, True)
    assert info["components"] == 3
    np.testing.assert_array_equal(F2, [[0, 1, 2, 3]])
    np.testing.assert_array_equal(V2, V[:4])
    # two disjoint quads of equal size: the lower one is kept, whichever comes first in V
    Vb = np.concatenate([V[:4] + np.float32(5), V[:4]])
    Fb = np.array([[4, 5, 6, 7], [0, 1, 2, 3]], np.int32)
    V3, F3, info = oracle.postprocess(Vb, Fb, False, True)
    assert info["components"] == 2
    np.testing.assert_array_equal(F3, [[0, 1, 2, 3]])
    np.testing.assert_array_equal(V3, Vb[4:8])


def _keys(V):
    V = V.astype(np.float32)
    ext = V.max(axis=0) - V.min(axis=0)
    sq = ext * ext
    diag = np.sqrt((sq[0] + sq[1]) + sq[2])
    eps = diag * np.float32(1e-5) if diag > 0 else np.float32(1e-5)
    inv = np.float32(1.0) / np.float32(eps)
    return _half_away(V * inv).astype(np.int64)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_weld_matches_brute_force_keys(seed):
    rng = np.random.default_rng(seed)
    base = rng.uniform(-3, 5, (300, 3)).astype(np.float32)
    # duplicates exact and within a quarter step; a spread that makes eps about 1e-4
    dup = base[rng.integers(0, 300, 200)]
    jit = dup + rng.uniform(-1, 1, dup.shape).astype(np.float32) * np.float32(1e-7)
    # pairs 3 steps apart in y only: distinct keys, so both stay (a key that dropped or mixed up
    # a coordinate, or a larger eps, would merge them)
    eps = np.float32(1e-5) * np.sqrt(np.float32(3 * 8.0 ** 2))
    apart = base[:40] + np.array([0, 3, 0], np.float32) * eps
    V = np.concatenate([base, dup, jit, apart])
    V = V[rng.permutation(V.shape[0])]
    F = rng.integers(0, V.shape[0], (400, 4)).astype(np.int32)
    F[:50, 2] = F[:50, 0]          # three distinct: kept
    F[50:80, 1] = F[50:80, 0]
    F[50:80, 3] = F[50:80, 2]      # two distinct: dropped
    K = _keys(V)
    _, first, inverse = np.unique(K, axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first)            # distinct keys in order of first vertex
    new = np.empty(order.size, np.int64)
    new[order] = np.arange(order.size)
    to = new[inverse.ravel()]
    W = V[np.sort(first)]
    G = to[F]
    distinct = np.array([len(set(r)) for r in G.tolist()])
    G = G[distinct >= 3]
    V2, F2 = oracle.weld(V, F)
    assert W.shape[0] == 340  # 300 base points and the 40 points apart from them
    np.testing.assert_array_equal(V2, W)
    np.testing.assert_array_equal(F2, G)
    assert (distinct[50:80] == 2).all()


def _blobs(n=40):
    """DC mesh of three disjoint spheres of radii 0.12, 0.2, 0.08 (oracle extraction)."""
    ax = (np.arange(n, dtype=np.float32) / np.float32(n - 1))
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    d = np.full(x.shape, np.inf, np.float32)
    for c, r in (((0.25, 0.25, 0.3), 0.12), ((0.65, 0.6, 0.6), 0.2), ((0.2, 0.75, 0.75), 0.08)):
        d = np.minimum(d, np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) - r)
    h = 1.0 / (n - 1)
    R = oracle.run(d.astype(np.float32), n, n, n, h, h, h)
    return R.V32, R.F


def test_floaters_keep_the_largest_component_scipy():
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    V, F = _blobs()
    Vw, Fw = oracle.weld(V, F)
    # faces joined by edges used by exactly two face sides
    sides = {}
    for q, f in enumerate(Fw.tolist()):
        for e in range(4):
            u, v = f[e], f[(e + 1) % 4]
            if u != v:
                sides.setdefault((min(u, v), max(u, v)), []).append(q)
    pairs = np.array([s for s in sides.values() if len(s) == 2])
    A = coo_matrix((np.ones(len(pairs)), (pairs[:, 0], pairs[:, 1])), shape=(len(Fw),) * 2)
    nc, lab = connected_components(A, directed=False)
    sizes = np.bincount(lab)
    best = np.flatnonzero(sizes == sizes.max())
    keep_lab = min(best, key=lambda c: np.flatnonzero(lab == c)[0])
    kept = Fw[lab == keep_lab]
    want_F, ids = _first_use(kept)
    V2, F2, info = oracle.postprocess(V, F, False, True)
    assert info["components"] == nc == 3
    np.testing.assert_array_equal(F2, want_F)
    np.testing.assert_array_equal(V2, Vw[ids])


def _frustum(n, r0=1.0, r1=2.0):
    ang = 2 * np.pi * np.arange(n) / n
    inner = np.stack([r0 * np.cos(ang), r0 * np.sin(ang), np.zeros(n)], 1)
    outer = np.stack([r1 * np.cos(ang), r1 * np.sin(ang), np.ones(n)], 1)
    V = np.concatenate([inner, outer]).astype(np.float32)
    F = np.array([[k, (k + 1) % n, n + (k + 1) % n, n + k] for k in range(n)], np.int32)
    return V, F


def test_ring_fill_of_regular_octagons_in_closed_form():
    # Loop of a regular n-gon of radius r: mean edge s = 2 r sin(pi/n); r > 1.25 s, so one ring at
    # radius r - s (each vertex stepped s towards the centroid), then r - s <= 1.25 s: a centre
    # fan. Per loop: n + 1 vertices, n ring quads and n fan triangles.
    n = 8
    V, F = _frustum(n)
    V2, F2, info = oracle.postprocess(V, F, True, False)
    assert info["loops_filled"] == 2
    assert info["added_vertices"] == 2 * (n + 1) and info["added_faces"] == 4 * n
    new = V2[2 * n:].astype(np.float64)
    rad = np.hypot(new[:, 0], new[:, 1])
    s0, s1 = 2 * np.sin(np.pi / n), 4 * np.sin(np.pi / n)
    lo = new[:, 2] < 0.5
    np.testing.assert_allclose(np.sort(rad[lo]), [0.0] + [1 - s0] * n, atol=2e-6)
    np.testing.assert_allclose(np.sort(rad[~lo]), [0.0] + [2 - s1] * n, atol=2e-6)
    _closed_and_oriented(V2, F2)


def test_loops_longer_than_4096_stay_open():
    V, F = _frustum(5000, 1.0, 1.5)
    V2, F2, info = oracle.postprocess(V, F, True, False)
    assert info["boundary_edges"] == 10000 and info["loops_filled"] == 0
    want_F, ids = _first_use(F)  # only renumbered by first use
    np.testing.assert_array_equal(F2, want_F)
    np.testing.assert_array_equal(V2, V[ids])


def _closed_and_oriented(V, F):
    d = _edges(F)
    assert len(set(d)) == len(d), "a directed side repeats (inconsistent orientation)"
    ds = set(d)
    assert all((v, u) in ds for u, v in d), "a side without its opposite (open boundary)"
    n_edges = len(d) // 2
    chi = V.shape[0] - n_edges + F.shape[0]
    return chi


@pytest.mark.parametrize("n", [16, 33])
def test_hole_closing_makes_a_cut_sphere_closed_genus_zero(n):
    ax = np.arange(n, dtype=np.float32) / np.float32(n - 1)
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    d = (np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2 + z ** 2) - 0.45).astype(np.float32)
    h = 1.0 / (n - 1)
    R = oracle.run(d, n, n, n, h, h, h)
    V2, F2, info = oracle.postprocess(R.V32, R.F, True, True)
    assert info["loops_filled"] == 1
    assert _closed_and_oriented(V2, F2) == 2
    # without hole closing the cut stays open: Euler characteristic of a disc
    V3, F3, _ = oracle.postprocess(R.V32, R.F, False, True)
    n_edges = len({(min(u, v), max(u, v)) for u, v in _edges(F3)})
    assert V3.shape[0] - n_edges + F3.shape[0] == 1
