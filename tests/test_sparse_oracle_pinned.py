"""Pins the oracle's restatement of the reference's sparse brick / multi-vertex backends
(oracle/sparse_oracle.cpp; src/cuda/dc_cuda_sparse.cu) against what the mathematics and the
already pinned dense oracle fix, on the CPU. The reference ships no test or fixture for these
backends (its src/cpp/dc_tests.cpp covers the CPU backend only), so every check here is
independent of the restatement:

- the 3x3 Jacobi pseudo-inverse against numpy's eigendecomposition (full rank, rank-deficient
  with the 0.01 cut, all values cut);
- hermite entries of fields whose gradients are known exactly: a linear field (trilinear and
  central-difference normals equal the plane normal), a quadratic (central differences are exact
  derivatives, so the edge-blended normal is the gradient at the point);
- vertices of a linear field lie on the plane wherever the solve was not clamped;
- topology of smooth grids equals the dense oracle's (pinned by the reference's own fixtures) as
  directed face cycles over cells, once the extra near-surface cells are mapped away;
- the activation rule (bricks with both signs; cells with both signs or a corner within 0.45 vx)
  brute-forced over every brick and cell;
- a sheet thinner than a cell: four thin z edges per cell, two clusters of opposite normals, both
  vertices at the sheet's centre, both windings of each quad.
tests/test_oracle_mutations.py plants mistakes in the oracle and checks these tests catch them.
"""
import numpy as np
import pytest

import grids
import oracle
from contouring_hip import synth


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


# ---------------------------------------------------------------------------------------------
# The Jacobi pseudo-inverse (dc_cuda_sparse.cu:178-269)
# ---------------------------------------------------------------------------------------------
def _pinv_ref(B, d, thr):
    w, V = np.linalg.eigh(np.asarray(B, np.float64))
    inv = np.where(w > thr, 1.0 / np.where(w > thr, w, 1.0), 0.0)
    return V @ (inv * (V.T @ np.asarray(d, np.float64)))


def test_jacobi_pinv_matches_eigendecomposition():
    rng = np.random.default_rng(1)
    for _ in range(200):
        N = rng.normal(size=(rng.integers(3, 9), 3))
        N /= np.linalg.norm(N, axis=1, keepdims=True)
        B = (N.T @ N).astype(np.float32)
        d = rng.normal(size=3).astype(np.float32)
        y, singular = oracle.sparse_jacobi(B, d)
        want = _pinv_ref(B, d, 0.01)
        assert not singular
        assert np.allclose(y, want, rtol=2e-4, atol=2e-5 * np.abs(want).max()), (B, d, y, want)


def test_jacobi_pinv_drops_small_singular_values():
    rng = np.random.default_rng(2)
    for rank in (1, 2):
        for _ in range(100):
            N = rng.normal(size=(rank, 3))
            Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
            # normals spanning `rank` directions only, plus a tiny component below the cut
            B = sum(np.outer(q, q) * s for q, s in zip(Q.T[:rank], rng.uniform(0.5, 3.0, rank)))
            B = B + np.outer(Q.T[2], Q.T[2]) * 0.004
            B = B.astype(np.float32)
            d = rng.normal(size=3).astype(np.float32)
            y, singular = oracle.sparse_jacobi(B, d)
            want = _pinv_ref(B, d, 0.01)
            assert not singular
            assert np.allclose(y, want, rtol=1e-3, atol=1e-4 * np.abs(want).max() + 1e-6)
    # every singular value at or under the cut: flagged, and nothing solved
    B = np.diag([0.01, 0.005, 0.0]).astype(np.float32)
    y, singular = oracle.sparse_jacobi(B, [1.0, 2.0, 3.0])
    assert singular and np.all(y == 0.0)
    # a diagonal system needs no rotation: exactly d / s
    y, singular = oracle.sparse_jacobi(np.diag([2.0, 4.0, 0.5]), [1.0, 1.0, 1.0])
    assert not singular and np.array_equal(y, np.float32([0.5, 0.25, 2.0]))


# ---------------------------------------------------------------------------------------------
# Hermite entries of fields with known gradients
# ---------------------------------------------------------------------------------------------
def _cells_with_entries(g, mode, limit=400):
    rng = np.random.default_rng(7)
    out = []
    for _ in range(limit):
        ijk = [int(rng.integers(1, n - 2)) for n in (g.nx, g.ny, g.nz)]
        P, N = oracle.sparse_cell(g.values, (g.nx, g.ny, g.nz), (g.vx, g.vy, g.vz),
                                  (g.ox, g.oy, g.oz), ijk, mode)
        if len(P):
            out.append((ijk, P, N))
    return out


def test_linear_field_entries_lie_on_the_plane_with_its_normal():
    a = np.float32([0.3, -0.5, 0.8])
    g = grids.tilted_plane(20, normal=tuple(a), d=0.1)
    nrm = _unit(a)
    for mode in (0, 1):
        cells = _cells_with_entries(g, mode)
        assert len(cells) > 20
        for ijk, P, N in cells:
            assert np.allclose(N, nrm, atol=2e-6), (mode, ijk, N)
            lin = P.astype(np.float64) @ a.astype(np.float64) - 0.1
            assert np.abs(lin).max() < 2e-6 * np.linalg.norm(a), (mode, ijk, lin)


def test_trilinear_normals_away_from_the_cell_centre():
    """The trilinear gradient's weights: on a linear field every entry's normal is exact however
    far from the cell centre the crossing is (a wrong weight only cancels at u = v = w)."""
    g = grids.tilted_plane(16, normal=(0.9, 0.2, -0.35), d=0.03)
    for ijk, P, N in _cells_with_entries(g, 0, 600):
        assert np.allclose(N, _unit([0.9, 0.2, -0.35]), atol=2e-6), (ijk, N)


def test_edge_gradient_of_a_quadratic_is_the_gradient_at_the_point():
    """f = |x - c|^2 - r^2: central differences at interior samples are the exact derivatives
    2(x - c), linear in x, so blending the ends' gradients at the crossing's t gives 2(p - c)."""
    n = 18
    h = np.float32(1.0) / np.float32(n - 1)
    ax = (np.arange(n, dtype=np.float32) * h).astype(np.float32)
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    c = np.float32([0.47, 0.52, 0.5])
    f = ((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 - np.float32(0.3) ** 2)
    g = grids._grid(f.astype(np.float32), (n, n, n), (h, h, h), (0.0, 0.0, 0.0), "quadratic")
    checked = 0
    for ijk, P, N in _cells_with_entries(g, 1, 800):
        for p, nn in zip(P, N):
            want = _unit(p.astype(np.float64) - c)
            assert np.allclose(nn, want, atol=3e-5), (ijk, p, nn, want)
            checked += 1
    assert checked > 50


def test_crossing_points_are_linear_zeros_of_their_edge():
    g = synth.csg8(24, seed=5)
    v = np.asarray(g.values, np.float32).reshape(g.nz, g.ny, g.nx)
    for ijk, P, _ in _cells_with_entries(g, 0, 300):
        i, j, k = ijk
        for p in P:
            # the point sits on one of the cell's 12 edges, at the linear zero of its values
            q = (p - np.float32([g.ox, g.oy, g.oz])) / np.float32(g.vx) - np.float32(ijk)
            free = [ax for ax in range(3) if 1e-4 < q[ax] < 1 - 1e-4]
            assert len(free) <= 1, (ijk, q)
            if not free:
                continue
            ax = free[0]
            lo = [int(round(t)) for t in q]
            lo[ax] = 0
            hi = list(lo)
            hi[ax] = 1
            fa = float(v[k + lo[2], j + lo[1], i + lo[0]])
            fb = float(v[k + hi[2], j + hi[1], i + hi[0]])
            if fa * fb <= 0:
                t = q[ax]
                assert abs((1 - t) * fa + t * fb) <= 1e-4 * (abs(fa) + abs(fb)), (ijk, q, fa, fb)


# ---------------------------------------------------------------------------------------------
# Vertices and topology
# ---------------------------------------------------------------------------------------------
def test_plane_vertices_on_the_plane_unless_clamped():
    a = np.float64([0.3, -0.5, 0.8])
    g = grids.tilted_plane(20, normal=tuple(a), d=0.1)
    for mv in (False, True):
        r = oracle.sparse_grid(g, multi_vertex=mv)
        V = r["vertices"].astype(np.float64)
        cells = r["active"]
        cx, cy = g.nx - 1, g.ny - 1
        ii, jj, kk = cells % cx, (cells // cx) % cy, cells // (cx * cy)
        lo = np.stack([g.ox + ii * g.vx, g.oy + jj * g.vy, g.oz + kk * g.vz], 1)
        inside = np.all((V > lo + 1e-6) & (V < lo + g.vx - 1e-6), axis=1)
        # cells active only for a corner near zero have no crossing: their vertex is the centre
        v = np.asarray(g.values, np.float32).reshape(g.nz, g.ny, g.nx)
        corners = np.stack([v[kk + dk, jj + dj, ii + di] for dk in (0, 1) for dj in (0, 1)
                            for di in (0, 1)], 1)
        crossing = (corners < 0).any(axis=1) & (corners > 0).any(axis=1)
        inside &= crossing
        on = np.abs(V @ a - 0.1) / np.linalg.norm(a)
        assert inside.sum() > 100
        assert on[inside].max() < 1e-5 * g.vx, on[inside].max()
        assert r["stats"]["two_vertex_cells"] == 0  # one normal direction: nothing to split


def _canon(cycle):
    c = [int(x) for x in cycle]
    k = c.index(min(c))
    return tuple(c[k:] + c[:k])


@pytest.mark.parametrize("make", [lambda: synth.sphere_unit(48),
                                  lambda: grids.ragged_sphere(30, 41, 23),
                                  lambda: grids.uneven_blob()],
                         ids=["sphere48", "ragged", "uneven_blob"])
def test_smooth_topology_equals_the_dense_oracle(make):
    """On smooth grids without exact zeros or thin edges every sparse quad is a dense quad with
    the same directed cycle of cells, and every dense active cell is a sparse active cell."""
    g = make()
    d = oracle.run(g.values, g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz)
    for mv in (False, True):
        s = oracle.sparse_grid(g, multi_vertex=mv)
        # vertex -> cell of the sparse mesh (one vertex per cell here)
        assert s["stats"]["two_vertex_cells"] == 0
        D = sorted(_canon(d.active[f]) for f in d.F)
        S = sorted(_canon(s["active"][f]) for f in s["faces"])
        assert D == S
        assert set(d.active.tolist()) <= set(s["active"].tolist())


def test_activation_rule_brute_force():
    for g, brick in ((grids.near_noise_grid(13, 9, neg=0.4), 3), (synth.csg8(40, seed=2), 8),
                     (grids.slab_grid(24), 5), (grids.zeros_grid(), 2), (grids.flat_zero_grid(16), 4)):
        v = np.asarray(g.values, np.float32).reshape(g.nz, g.ny, g.nx)
        cx, cy, cz = g.nx - 1, g.ny - 1, g.nz - 1
        near = np.float32(0.45) * np.float32(g.vx)
        want = []
        bricks = 0
        nb = [(c + brick - 1) // brick for c in (cx, cy, cz)]
        active_brick = np.zeros(nb[::-1], bool)
        for bz in range(nb[2]):
            for by in range(nb[1]):
                for bx in range(nb[0]):
                    s = v[bz * brick:bz * brick + brick + 1, by * brick:by * brick + brick + 1,
                          bx * brick:bx * brick + brick + 1]
                    active_brick[bz, by, bx] = (s < 0).any() and (s > 0).any()
                    bricks += active_brick[bz, by, bx]
        for k in range(cz):
            for j in range(cy):
                for i in range(cx):
                    if not active_brick[k // brick, j // brick, i // brick]:
                        continue
                    c = v[k:k + 2, j:j + 2, i:i + 2]
                    if ((c < 0).any() and (c > 0).any()) or (np.abs(c) <= near).any():
                        want.append(i + cx * (j + cy * k))
        r = oracle.sparse_grid(g, brick_size=brick)
        assert r["active"].tolist() == want, g.name
        assert r["stats"]["active_bricks"] == bricks
        assert r["stats"]["total_bricks"] == nb[0] * nb[1] * nb[2]


def test_thin_sheet_in_closed_form():
    """slab_grid: scale 2 (|z - c| - 0.3 h) with c halfway between sample layers k0 and k0 + 1.
    Every cell of layer k0 has four thin z edges (both ends 0.4 h, gradients (0,0,-2) and
    (0,0,+2)) and no other entry: eight entries at the sheet's centre plane, normals -z and +z
    alternating. One brick covers the grid."""
    n = 24
    g = grids.slab_grid(n)
    h = np.float64(g.vx)
    k0 = n // 2
    c = (np.float32(k0) * np.float32(h) + np.float32(0.5) * np.float32(h))
    cx = cy = n - 1
    for mv in (False, True):
        r = oracle.sparse_grid(g, brick_size=100, multi_vertex=mv)
        act, vc, V, Nn = r["active"], r["vcount"], r["vertices"], r["vnormals"]
        first = np.concatenate([[0], np.cumsum(vc)[:-1]])
        ii, jj, kk = act % cx, (act // cx) % cy, act // (cx * cy)
        # layer k0 away from the ball (x, y > 0.45) and the border
        sel = (kk == k0) & (ii * h > 0.45) & (jj * h > 0.45) & (ii < cx - 1) & (jj < cy - 1)
        assert sel.sum() > 50
        for q in np.flatnonzero(sel):
            want = np.float64([g.ox + (ii[q] + 0.5) * h, g.oy + (jj[q] + 0.5) * h, c])
            nv = 2 if mv else 1
            assert vc[q] == nv
            for t in range(nv):
                assert np.allclose(V[first[q] + t], want, atol=1e-6), (q, V[first[q] + t], want)
            if mv:
                assert np.array_equal(Nn[first[q]], np.float32([0, 0, -1]))
                assert np.array_equal(Nn[first[q] + 1], np.float32([0, 0, 1]))
        # the sheet's z edges emit both windings: each such quad is followed by its reverse
        F = r["faces"]
        rev = np.all(F[1:] == F[:-1][:, [0, 3, 2, 1]], axis=1)
        assert rev.sum() > 100


def test_thin_edges_need_opposed_gradients():
    """A true distance sheet (slope 1) thinner than a cell has end gradients of +-0.5 (the
    central differences straddle the kink): their dot, -0.25, is above -0.3, so no edge is thin
    and the sheet disappears."""
    g = grids.slab_grid(24, scale=1.0, half=0.05)
    r = oracle.sparse_grid(g, brick_size=100, multi_vertex=True)
    assert r["stats"]["two_vertex_cells"] == 0
    s = grids.slab_grid(24, scale=2.0)
    assert oracle.sparse_grid(s, brick_size=100, multi_vertex=True)["stats"]["two_vertex_cells"] > 0


def test_bad_arguments():
    g = synth.sphere_unit(8)
    for kw in ({"brick_size": 0}, {"normal_mode": 3}, {"normal_mode": -1}):
        with pytest.raises(ValueError):
            oracle.sparse_grid(g, **kw)
