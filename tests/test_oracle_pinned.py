"""The oracle pinned by mathematics and by hand-built grids with expected values, independently of
its own formulas (CPU only). Each test names the oracle mistake it catches;
tests/test_oracle_mutations.py plants those mistakes in a copy of the oracle and checks that the
named test fails.

- trilinear field: on a grid sampling f = a + bx + cy + dz + exy + fyz + gzx + hxyz exactly (dyadic
  coefficients and spacing, so FP32 samples are exact), each cell's trilinear interpolant is f, so
  every raw hermite point is f's root on its edge and every raw normal is grad f / |grad f| there
  (Cell.cpp:80-155, 162-212). Kills a wrong weight or term in the trilinear gradient.
- averaging selection (hermite_update.cpp:171-438): hand-built 5^3 grids where the 4 cells around
  an edge have mirror-image raw normals, so an averaged edge ends with a normal whose cross
  components cancel and an unaveraged one keeps them: an edge with one endpoint on the border and
  one strictly inside is averaged; an edge with both endpoints on the border is not; an edge with a
  zero endpoint is averaged only when all 4 cells are active. Kills "both endpoints inside" and
  "average over the active cells only".
- quad order and orientation (contouring.cpp:141-207, 607-729): quads enumerated by brute force in
  the reference's loop order from the signs alone, and each quad's orientation recomputed from the
  oracle's FP64 vertices and an independent NumPy central-difference gradient with clamped
  neighbours. Kills a changed face order and a changed border clamp of gradientAt.
"""
import numpy as np
import pytest

import oracle

# Cell corner offsets (contouring.cpp:232-241) and edges as corner pairs (include/Cell.h:26-30).
CORNERS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
EDGES = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6),
         (3, 7)]


def _run(vals_zyx, h=0.25):
    nz, ny, nx = vals_zyx.shape
    return oracle.run(np.ascontiguousarray(vals_zyx, np.float32).ravel(), nx, ny, nz, h, h, h)


def _cell_edge(cell, p0, p1):
    """Local edge index of the grid edge p0-p1 in cell (i, j, k)."""
    d0 = tuple(p0[q] - cell[q] for q in range(3))
    d1 = tuple(p1[q] - cell[q] for q in range(3))
    for e, (a, b) in enumerate(EDGES):
        if (CORNERS[a], CORNERS[b]) in ((d0, d1), (d1, d0)):
            return e
    raise ValueError("edge not in cell")


def _slots(res, dims, p0, p1):
    """(vertex id, local edge) of every active cell containing the grid edge p0-p1."""
    nx, ny, nz = dims
    cx, cy = nx - 1, ny - 1
    vid = {int(c): v for v, c in enumerate(res.active)}
    axis = [q for q in range(3) if p0[q] != p1[q]][0]
    out = []
    others = [q for q in range(3) if q != axis]
    for m1 in (0, 1):
        for m2 in (0, 1):
            cell = list(p0)
            cell[others[0]] -= m1
            cell[others[1]] -= m2
            if min(cell) < 0 or cell[0] >= nx - 1 or cell[1] >= ny - 1 or cell[2] >= nz - 1:
                continue
            lin = cell[0] + cx * (cell[1] + cy * cell[2])
            if lin in vid:
                out.append((vid[lin], _cell_edge(cell, p0, p1)))
    return out


# ---------------------------------------------------------------------------------------------
# trilinear field
# ---------------------------------------------------------------------------------------------
COEF = dict(a=-1.0, b=0.5, c=0.25, d=0.75, e=0.125, f=-0.25, g=0.375, h=0.5)


def _tri(x, y, z):
    k = COEF
    return (k["a"] + k["b"] * x + k["c"] * y + k["d"] * z + k["e"] * x * y + k["f"] * y * z
            + k["g"] * z * x + k["h"] * x * y * z)


def _tri_grad(p):
    k = COEF
    x, y, z = p
    return np.array([k["b"] + k["e"] * y + k["g"] * z + k["h"] * y * z,
                     k["c"] + k["e"] * x + k["f"] * z + k["h"] * x * z,
                     k["d"] + k["f"] * y + k["g"] * x + k["h"] * x * y])


def test_trilinear_field_hermite_points_and_normals_in_closed_form():
    """Mutation: a wrong weight or term in the trilinear gradient (Cell.cpp:80-155)."""
    n, hs = 9, 0.25
    ax = np.arange(n) * hs
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    vals = _tri(x, y, z)
    assert np.array_equal(vals.astype(np.float32).astype(np.float64), vals)  # samples are exact
    res = _run(vals, hs)
    assert res.n_vertices > 20
    mask, pts, nrm = res.hermite(averaged=False)
    checked = 0
    for v in range(res.n_vertices):
        for e in range(12):
            if not (mask[v] >> e) & 1:
                continue
            p = pts[v, e]
            assert abs(_tri(*p)) <= 1e-14, (v, e, _tri(*p))  # the root of f on the edge
            gr = _tri_grad(p)
            want = gr / np.linalg.norm(gr)
            np.testing.assert_allclose(nrm[v, e], want, rtol=0, atol=1e-13)
            checked += 1
    assert checked > 100
    # every cell's interpolant is the same f: averaging changes no normal
    _, _, avg = res.hermite(averaged=True)
    m = np.array([[(mask[v] >> e) & 1 for e in range(12)] for v in range(res.n_vertices)], bool)
    np.testing.assert_allclose(avg[m], nrm[m], rtol=0, atol=1e-13)


# ---------------------------------------------------------------------------------------------
# averaging selection on hand-built 5^3 grids (samples indexed [k][j][i])
# ---------------------------------------------------------------------------------------------
def _normals_at(res, dims, p0, p1, averaged):
    _, _, nrm = res.hermite(averaged=averaged)
    return [nrm[v, e] for v, e in _slots(res, dims, p0, p1)]


def _assert_averaged(res, dims, p0, p1):
    raw = _normals_at(res, dims, p0, p1, False)
    avg = _normals_at(res, dims, p0, p1, True)
    assert len(raw) == 4
    s = np.sum([r for r in raw if np.linalg.norm(r) > 0], axis=0)
    want = s / np.linalg.norm(s)
    for a in avg:
        np.testing.assert_allclose(a, want, rtol=0, atol=1e-12)
    # the mirror-image cells cancel the cross components: the result is the edge direction
    axis = [q for q in range(3) if p0[q] != p1[q]][0]
    others = [q for q in range(3) if q != axis]
    assert np.abs(want[others]).max() <= 1e-12
    assert min(np.abs(r[others]).max() for r in raw) > 1e-3  # each raw normal had them


def _assert_raw_kept(res, dims, p0, p1, n_cells):
    raw = _normals_at(res, dims, p0, p1, False)
    avg = _normals_at(res, dims, p0, p1, True)
    assert len(raw) == n_cells
    for r, a in zip(raw, avg):
        np.testing.assert_array_equal(a, r)
    if n_cells > 1:  # the cells' raw normals differ, so an average would have shown
        assert max(np.abs(raw[0] - r).max() for r in raw[1:]) > 1e-3


def test_edge_with_one_border_endpoint_is_averaged():
    """Mutation: averaging only edges whose two endpoints are both strictly inside (the reference
    needs one, hermite_update.cpp:190-195)."""
    v = np.ones((5, 5, 5))
    v[2, 2, 0] = -1.0          # sample (0, 2, 2) on the x = 0 border
    v[2, 2, 4] = -1.0          # sample (4, 2, 2) on the x = 4 border
    res = _run(v)
    dims = (5, 5, 5)
    _assert_averaged(res, dims, (0, 2, 2), (1, 2, 2))  # lower endpoint on the border
    _assert_averaged(res, dims, (3, 2, 2), (4, 2, 2))  # upper endpoint on the border


def test_edge_with_both_endpoints_on_the_border_is_not_averaged():
    """The border test's other half (hermite_update.cpp:190-195): an edge with neither endpoint
    strictly inside keeps its cells' raw normals. (Dropping the test outright makes the oracle read
    cells at index -1; it is not a mutation this suite can fail cleanly.)"""
    v = np.ones((5, 5, 5))
    v[2, 0, 2] = -1.0          # sample (2, 0, 2) on the y = 0 border
    res = _run(v)
    dims = (5, 5, 5)
    # x edge (1,0,2)-(2,0,2): both endpoints have y = 0; its two cells keep their raw normals
    _assert_raw_kept(res, dims, (1, 0, 2), (2, 0, 2), 2)
    # the z edge (2,0,1)-(2,0,2) likewise (both at y = 0)
    _assert_raw_kept(res, dims, (2, 0, 1), (2, 0, 2), 2)


def test_zero_endpoint_edge_averaged_only_with_four_active_cells():
    """Mutation: averaging over the active cells only, when some of the 4 are inactive
    (hermite_update.cpp:298-303 requires all four)."""
    dims = (5, 5, 5)
    p0, p1 = (1, 2, 2), (2, 2, 2)   # x edge, S(p0) = 0, S(p1) > 0: a hermite sample, no quad
    v = np.ones((5, 5, 5))
    v[2, 2, 1] = 0.0
    # negative corners in three of the four cells around the edge: (1,2,2) via (1,3,3),
    # (1,1,2) via (1,1,3), (1,2,1) via (1,3,1); cell (1,1,1) has no negative corner
    for (i, j, k) in ((1, 3, 3), (1, 1, 3), (1, 3, 1)):
        v[k, j, i] = -1.0
    res = _run(v)
    _assert_raw_kept(res, dims, p0, p1, 3)
    # with a negative corner in the fourth cell too, the edge is averaged
    v[1, 1, 1] = -1.0               # sample (1, 1, 1)
    res = _run(v)
    _assert_averaged(res, dims, p0, p1)


# ---------------------------------------------------------------------------------------------
# quad order and orientation by brute force
# ---------------------------------------------------------------------------------------------
def _gradient_at(S, coords, i, j, k):
    """contouring.cpp:141-207 restated: central differences, neighbours clamped to the grid,
    normalized; coords[q] are the grid's coordinate arrays."""
    nz, ny, nx = S.shape
    idx = [i, j, k]
    size = [nx, ny, nz]
    g = np.zeros(3)
    for q in range(3):
        lo, hi = list(idx), list(idx)
        lo[q] = max(idx[q] - 1, 0)
        hi[q] = min(idx[q] + 1, size[q] - 1)
        d = coords[q][hi[q]] - coords[q][lo[q]]
        if abs(d) > 1e-12:
            g[q] = (S[hi[2], hi[1], hi[0]] - S[lo[2], lo[1], lo[0]]) / d
    nrm = np.linalg.norm(g)
    return g / nrm if nrm > 0 else g


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_quad_order_and_orientation_brute_force(seed):
    """Mutations: a changed quad enumeration order, a changed border clamp in gradientAt."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = 8, 7, 6
    S = rng.normal(size=(nz, ny, nx)).astype(np.float32).astype(np.float64)
    S[rng.random(S.shape) < 0.05] = 0.0   # some exact zeros
    res = oracle.run(S.astype(np.float32).ravel(), nx, ny, nz, 0.25, 0.5, 0.375, 0.0, -1.0, 0.5)
    cx, cy, cz = nx - 1, ny - 1, nz - 1
    vid = {int(c): v for v, c in enumerate(res.active)}

    def cell(i, j, k):
        return vid.get(i + cx * (j + cy * k), -1)

    coords = [np.float32(0.0) + np.arange(nx, dtype=np.float32) * np.float32(0.25),
              np.float32(-1.0) + np.arange(ny, dtype=np.float32) * np.float32(0.5),
              np.float32(0.5) + np.arange(nz, dtype=np.float32) * np.float32(0.375)]
    coords = [c.astype(np.float64) for c in coords]
    expect = []  # (c0, c1, c2, c3, lower endpoint, upper endpoint)
    for i in range(cx):
        for j in range(1, cy):
            for k in range(1, cz):
                if S[k, j, i] * S[k, j, i + 1] < 0:
                    expect.append((cell(i, j - 1, k - 1), cell(i, j, k - 1), cell(i, j, k),
                                   cell(i, j - 1, k), (i, j, k), (i + 1, j, k)))
    for i in range(1, cx):
        for j in range(cy):
            for k in range(1, cz):
                if S[k, j, i] * S[k, j + 1, i] < 0:
                    expect.append((cell(i - 1, j, k - 1), cell(i, j, k - 1), cell(i, j, k),
                                   cell(i - 1, j, k), (i, j, k), (i, j + 1, k)))
    for i in range(1, cx):
        for j in range(1, cy):
            for k in range(cz):
                if S[k, j, i] * S[k + 1, j, i] < 0:
                    expect.append((cell(i - 1, j - 1, k), cell(i, j - 1, k), cell(i, j, k),
                                   cell(i - 1, j, k), (i, j, k), (i, j, k + 1)))
    assert len(expect) == res.n_faces > 20
    V = res.V64
    for f, (c0, c1, c2, c3, e0, e1) in enumerate(expect):
        assert min(c0, c1, c2, c3) >= 0
        gn = _gradient_at(S, coords, *e0) + _gradient_at(S, coords, *e1)
        pa, pb, pc, pd = V[c0], V[c1], V[c2], V[c3]
        qn = np.cross(pb - pa, pd - pa)
        if qn @ qn < 1e-24:
            qn = np.cross(pc - pa, pd - pa)
        dot = qn @ gn if (gn @ gn > 0 and qn @ qn > 0) else 0.0
        assert abs(res.dots[f] - dot) <= 1e-12 * max(1.0, abs(dot)), (f, res.dots[f], dot)
        want = [c0, c3, c2, c1] if dot < 0 else [c0, c1, c2, c3]
        assert list(res.F[f]) == want, (f, list(res.F[f]), want)
