"""The mesh-to-SDF oracle (oracle/m2s_oracle.cpp, the restatement of the reference's
compute_mesh_sdf_device_cuda, src/cuda/mesh_to_sdf_cuda.cu:740-973) pinned by what the
mathematics fixes, each test independent of the oracle's code (CPU):

- closest point on a triangle: an FP64 projection-or-edges routine (no Voronoi regions), every
  region hit;
- distances of band samples of closed convex meshes (box, icosphere): the true FP64 distance to
  the mesh, and their signs: inside tests by face planes;
- signs on nonconvex closed meshes (torus, hollow box) and a sharp wedge: inside tests by ray
  parity;
- the weighted normal on a grid of one bin (each triangle visited once): at a vertex the
  angle-weighted pseudo-normal, on an edge the area-weighted sum of its two faces, inside a face
  its normal; signs of bumpy closed meshes there all agree with ray parity;
- the bins: which samples see a lone triangle follows the documented rule (n/8 bins per axis,
  boxes grown by 2 max(v), the 3x3x3 neighbourhood, the 2.5-bin band), and a degenerate triangle
  is seen by none;
- the flood: far exterior samples are +10, far samples deep inside a thick solid -10, those of a
  closed air cavity +10 (its band is positive), an empty mesh +10 everywhere;
- the relaxation: a column above a flat face against a 1-D Jacobi recurrence (4 passes, samples
  within 1.5 h kept);
- the voxelize option: solid voxels contain every voxel well inside the mesh and no voxel well
  outside, a closed cavity is filled, far samples take the solid's sign, and the one-voxel growth
  seals a hole two voxels across.

Each test names the planted oracle mistake it catches (tests/test_oracle_mutations.py).
"""
import numpy as np
import pytest

import oracle
from contouring_hip import synth

FAR = np.finfo(np.float32).max


def _unit(n):
    h = 1.0 / (n - 1)
    return (n, n, n), (0.0, 0.0, 0.0), (h, h, h)


def _grid_points(dims, origin, spacing):
    nx, ny, nz = dims
    xs = (np.float32(origin[0]) + np.arange(nx, dtype=np.float32) * np.float32(spacing[0]))
    ys = (np.float32(origin[1]) + np.arange(ny, dtype=np.float32) * np.float32(spacing[1]))
    zs = (np.float32(origin[2]) + np.arange(nz, dtype=np.float32) * np.float32(spacing[2]))
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    return np.stack([X, Y, Z], -1).astype(np.float64)  # (nz, ny, nx, 3)


def _seg_dist2(P, a, b):
    ab = b - a
    L = float(ab @ ab)
    t = np.zeros(P.shape[:-1]) if L == 0 else np.clip(((P - a) @ ab) / L, 0.0, 1.0)
    Q = a + t[..., None] * ab
    return ((P - Q) ** 2).sum(-1), Q


def _tri_dist2(P, a, b, c):
    """FP64 squared distance from points P (..., 3) to triangle (a, b, c) and the closest points:
    the projection on the plane when it falls inside, else the nearest of the three edges."""
    a, b, c = (np.asarray(x, np.float64) for x in (a, b, c))
    n = np.cross(b - a, c - a)
    nn = float(n @ n)
    best, Q = _seg_dist2(P, a, b)
    for u, v in ((b, c), (c, a)):
        d, q = _seg_dist2(P, u, v)
        m = d < best
        best = np.where(m, d, best)
        Q = np.where(m[..., None], q, Q)
    if nn == 0.0:
        return best, Q
    proj = P - (((P - a) @ n) / nn)[..., None] * n
    # barycentric coordinates of the projection
    v0, v1 = b - a, c - a
    d00, d01, d11 = v0 @ v0, v0 @ v1, v1 @ v1
    w = proj - a
    d20, d21 = w @ v0, w @ v1
    den = d00 * d11 - d01 * d01
    s = (d11 * d20 - d01 * d21) / den
    t = (d00 * d21 - d01 * d20) / den
    inside = (s >= 0) & (t >= 0) & (s + t <= 1)
    dp = ((P - proj) ** 2).sum(-1)
    best = np.where(inside, dp, best)
    Q = np.where(inside[..., None], proj, Q)
    return best, Q


def _mesh_dist(P, V, F):
    V = V.astype(np.float64)
    best = np.full(P.shape[:-1], np.inf)
    for f in F:
        d, _ = _tri_dist2(P, V[f[0]], V[f[1]], V[f[2]])
        best = np.minimum(best, d)
    return np.sqrt(best)


def _box_surface_dist(P, lo, hi):
    """FP64 distance from P to the surface of the axis-aligned box [lo, hi]."""
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    c, e = (lo + hi) / 2, (hi - lo) / 2
    q = np.abs(P - c) - e
    outside = np.sqrt((np.maximum(q, 0.0) ** 2).sum(-1))
    inside = np.minimum(q.max(-1), 0.0)
    return np.abs(outside + inside)


def _inside_convex(P, V, F):
    V = V.astype(np.float64)
    inside = np.ones(P.shape[:-1], bool)
    for f in F:
        a, b, c = V[f]
        n = np.cross(b - a, c - a)
        inside &= (P - a) @ n < 0
    return inside


def _inside_parity(P, V, F, direction=(0.5773, 0.5801, 0.5745)):
    """Ray parity along a fixed skew direction (Moller-Trumbore, FP64)."""
    V = V.astype(np.float64)
    d = np.asarray(direction, np.float64)
    d /= np.linalg.norm(d)
    hits = np.zeros(P.shape[:-1], np.int32)
    for f in F:
        a, b, c = V[f]
        e1, e2 = b - a, c - a
        h = np.cross(d, e2)
        det = e1 @ h
        if abs(det) < 1e-15:
            continue
        s = P - a
        u = (s @ h) / det
        q = np.cross(s, e1)
        v = (q @ d) / det
        t = (q @ e2) / det
        hits += ((u >= 0) & (v >= 0) & (u + v <= 1) & (t > 0)).astype(np.int32)
    return hits % 2 == 1


def _bins(dims, spacing):
    g = [max(1, n // 8) for n in dims]
    bd = [np.float32(np.float32(n) * np.float32(v)) / np.float32(gg)
          for n, v, gg in zip(dims, spacing, g)]
    return g, [np.float32(b) for b in bd]


# ---------------------------------------------------------------------------------------------


def test_point_triangle_closest_point():
    """Catches: a wrong region test in the closest-point routine (mutation corner_b_region)."""
    rng = np.random.default_rng(5)
    regions = set()
    for _ in range(400):
        a, b, c = rng.uniform(-1, 1, (3, 3)).astype(np.float32)
        n = np.cross(b - a, c - a)
        if np.linalg.norm(n) < 0.2:  # well-shaped triangles only
            continue
        for p in rng.uniform(-2, 2, (8, 3)).astype(np.float32):
            d, q = oracle.point_triangle(p, a, b, c)
            dr, qr = _tri_dist2(p.astype(np.float64), a, b, c)
            assert abs(d - dr) <= 1e-5 * dr + 1e-6, (p, a, b, c, d, dr)
            assert np.abs(q - qr).max() <= 1e-4, (q, qr)
            bary = np.linalg.lstsq(np.stack([b - a, c - a], 1).astype(np.float64),
                                   qr - a, rcond=None)[0]
            regions.add((bool(bary[0] > 1e-6), bool(bary[1] > 1e-6),
                         bool(bary.sum() < 1 - 1e-6)))
    assert len(regions) == 7  # three corners, three edges, the interior


@pytest.mark.parametrize("name", ["box", "icosphere"])
def test_convex_band_distances_and_signs(name):
    """Catches: a wrong sign rule or normal (mutation normal_flipped) and wrong distances."""
    V, F = synth.box_mesh() if name == "box" else synth.icosphere(2, radius=0.3)
    dims, origin, spacing = _unit(36)
    raw = oracle.mesh_to_sdf(V, F, dims, origin, spacing, raw=True)
    P = _grid_points(dims, origin, spacing)
    dt = _mesh_dist(P, V, F)
    inside = _inside_convex(P, V, F)
    _, bd = _bins(dims, spacing)
    near = dt < min(bd)
    assert (raw[near] != FAR).all()
    v = raw[near]
    assert np.all(np.abs(np.abs(v) - dt[near]) <= 1e-5 * dt[near] + 1e-6)
    clear = near & (dt > 1e-5)
    assert np.array_equal(raw[clear] < 0, inside[clear])
    # farther than the 2.5-bin band: never a distance
    assert (raw[dt > 2.5 * min(bd) * (1 + 1e-4)] == FAR).all()


@pytest.mark.parametrize("name", ["torus", "hollow_box"])
def test_nonconvex_band_signs(name):
    """Catches: signs from the last nearest triangle only (mutation no_tie_sum)."""
    V, F = synth.torus_mesh(nu=32, nv=16) if name == "torus" else synth.hollow_box_mesh()
    dims, origin, spacing = _unit(30)
    raw = oracle.mesh_to_sdf(V, F, dims, origin, spacing, raw=True)
    P = _grid_points(dims, origin, spacing)
    dt = _mesh_dist(P, V, F)
    inside = _inside_parity(P, V, F)
    ok = (raw != FAR) & (dt > 1e-4)
    assert ok.sum() > 1000
    assert np.array_equal(raw[ok] < 0, inside[ok])


def _wedge(angle_deg=20.0, length=0.6):
    """A closed triangular prism whose cross-section is an isosceles triangle with a sharp apex
    (the apex edge runs along x): two equal long faces meet at the sharp edge."""
    h = 0.3
    half = h * np.tan(np.radians(angle_deg) / 2)
    x0, x1 = 0.5 - length / 2, 0.5 + length / 2
    # cross-section in (y, z): apex at (0.5, 0.65), base at z = 0.35
    apex, bl, br = (0.5, 0.35 + h), (0.5 - half, 0.35), (0.5 + half, 0.35)
    V = np.array([[x0, *apex], [x0, *bl], [x0, *br], [x1, *apex], [x1, *bl], [x1, *br]],
                 np.float32)
    F = np.array([[0, 2, 1], [3, 4, 5],            # end caps (outward)
                  [0, 3, 5], [0, 5, 2],            # right long face
                  [0, 1, 4], [0, 4, 3],            # left long face
                  [1, 2, 5], [1, 5, 4]], np.int32)  # base
    return V, F


def test_sharp_edge_signs():
    """Outside a 20-degree edge each long face's own normal gives the wrong sign for part of the
    edge's region; the sum of both (the tie rule) gives the right one."""
    V, F = _wedge()
    dims, origin, spacing = _unit(48)
    raw = oracle.mesh_to_sdf(V, F, dims, origin, spacing, raw=True)
    P = _grid_points(dims, origin, spacing)
    dt = _mesh_dist(P, V, F)
    inside = _inside_convex(P, V, F)
    ok = (raw != FAR) & (dt > 1e-4)
    above = ok & (P[..., 2] > 0.66)  # samples over the sharp edge
    assert above.sum() > 100
    assert np.array_equal(raw[ok] < 0, inside[ok])


def _bumpy_icosphere(seed):
    V, F = synth.icosphere(1, radius=0.3)
    rng = np.random.default_rng(seed)
    c = np.array([0.5, 0.5, 0.5], np.float32)
    return (c + (V - c) * (1 + rng.uniform(-0.35, 0.35, (len(V), 1)))).astype(np.float32), F


def test_weighted_normals_in_one_bin():
    """On a grid of one bin every triangle is visited once, so a sample's summed normal is fixed
    by the closest feature: at a vertex, the angle-weighted pseudo-normal sum(angle_i n_i) over
    the triangles at that vertex (Baerentzen and Aanaes, 2005); on an edge, sum(2 area_i n_i) over
    its two triangles (the reference weights edges by area, :296-304); inside a face, that face's
    normal. The features are found independently in FP64. Catches: mutations no_tie_sum and
    angle_weight."""
    V, F = _bumpy_icosphere(2)
    n = 15
    dims, origin, spacing = (n, n, n), (0.12, 0.13, 0.11), (0.055, 0.054, 0.056)
    raw = oracle.mesh_to_sdf(V, F, dims, origin, spacing, raw=True)
    P = _grid_points(dims, origin, spacing)
    _, bd = _bins(dims, spacing)
    eps = 2e-3 * float(min(bd))
    Vd = V.astype(np.float64)
    normals, area2, dists, closest = [], [], [], []
    for f in F:
        a, b, c = Vd[f]
        nrm = np.cross(b - a, c - a)
        area2.append(np.linalg.norm(nrm))
        normals.append(nrm / area2[-1])
        d, q = _tri_dist2(P, a, b, c)
        dists.append(d)
        closest.append(q)
    dists, closest = np.stack(dists), np.stack(closest)
    counts = {"vertex": 0, "edge": 0, "face": 0}
    for z, y, x in np.argwhere(raw != FAR):
        d = dists[:, z, y, x]
        m = d.min()
        if m < 1e-8:
            continue
        tied = np.nonzero(d <= m * (1 + 1e-7) + 1e-15)[0]
        if (d <= m * (1 + 3e-5) + 1e-14).sum() != len(tied):
            continue  # a near-tie the reference's 1e-5 tolerance may or may not join
        q = closest[tied[0], z, y, x]
        at = [k for k in range(len(V)) if np.linalg.norm(q - Vd[k]) < eps]
        near = [k for k in range(len(V)) if np.linalg.norm(q - Vd[k]) < 10 * eps]
        if len(at) == 1 and np.linalg.norm(q - Vd[at[0]]) < 1e-9:
            kind, v = "vertex", at[0]
            inc = [i for i in range(len(F)) if v in F[i]]
            if sorted(inc) != sorted(tied.tolist()):
                continue
            want = np.zeros(3)
            for i in inc:
                j = list(F[i]).index(v)
                e1 = Vd[F[i][(j + 1) % 3]] - Vd[v]
                e2 = Vd[F[i][(j + 2) % 3]] - Vd[v]
                ang = np.arccos(np.clip(e1 @ e2 / np.linalg.norm(e1) / np.linalg.norm(e2), -1, 1))
                want += ang * normals[i]
        elif near:
            continue  # within the reference's corner radius of a vertex but not at it
        elif len(tied) == 2:
            kind = "edge"
            want = area2[tied[0]] * normals[tied[0]] + area2[tied[1]] * normals[tied[1]]
        elif len(tied) == 1:
            kind = "face"
            want = normals[tied[0]]
        else:
            continue
        _, _, got = oracle.m2s_sample_state(V, F, dims, origin, spacing, x, y, z)
        cos = float(got @ want) / (np.linalg.norm(got) * np.linalg.norm(want))
        assert cos > 1 - 1e-6, (kind, (x, y, z), got, want)
        counts[kind] += 1
    assert counts["vertex"] > 20 and counts["edge"] > 50 and counts["face"] > 100, counts


def test_single_bin_signs_match_ray_parity():
    """With every triangle visited once, the signs of a bumpy closed mesh's band samples all agree
    with ray parity. Catches: mutation no_tie_sum."""
    for seed in (2, 12):
        V, F = _bumpy_icosphere(seed)
        n = 15
        dims, origin, spacing = (n, n, n), (0.12, 0.13, 0.11), (0.055, 0.054, 0.056)
        raw = oracle.mesh_to_sdf(V, F, dims, origin, spacing, raw=True)
        P = _grid_points(dims, origin, spacing)
        dt = _mesh_dist(P, V, F)
        inside = _inside_parity(P, V, F)
        ok = (raw != FAR) & (dt > 1e-4)
        assert ok.sum() > 3000
        assert np.array_equal(raw[ok] < 0, inside[ok]), seed


def test_bins_follow_the_documented_rule():
    """Which samples see a lone triangle: catches mutations bin_size, margin and band."""
    dims, origin, spacing = (64, 64, 64), (0.0, 0.0, 0.0), (1 / 63, 1 / 63, 1 / 63)
    V = np.array([[0.40, 0.45, 0.50], [0.47, 0.46, 0.52], [0.43, 0.52, 0.49]], np.float32)
    F = np.array([[0, 1, 2]], np.int32)
    raw = oracle.mesh_to_sdf(V, F, dims, origin, spacing, raw=True)
    g, bd = _bins(dims, spacing)
    margin = np.float32(2.0) * np.float32(max(spacing))
    P = _grid_points(dims, origin, spacing).astype(np.float32)
    seen = np.ones(raw.shape, bool)
    for a in range(3):
        lo = np.float32(V[:, a].min() - margin)
        hi = np.float32(V[:, a].max() + margin)
        b0 = min(max(int(np.floor((lo - np.float32(origin[a])) / bd[a])), 0), g[a] - 1)
        b1 = min(max(int(np.floor((hi - np.float32(origin[a])) / bd[a])), 0), g[a] - 1)
        bins = np.clip(np.floor((P[..., a] - np.float32(origin[a])) / bd[a]), 0, g[a] - 1)
        seen &= (bins >= b0 - 1) & (bins <= b1 + 1)
    dt = _mesh_dist(P.astype(np.float64), V, F)
    lim = 2.5 * float(min(bd))
    clear = np.abs(dt - lim) > 1e-4 * lim
    want = seen & (dt <= lim)
    assert np.array_equal((raw != FAR)[clear], want[clear])
    assert want.sum() > 500 and (seen & ~want).sum() > 0  # both the band and the bins decide
    # a degenerate (collinear) triangle has no normal and is skipped
    V2 = np.array([[0.4, 0.4, 0.4], [0.5, 0.5, 0.5], [0.6, 0.6, 0.6]], np.float32)
    assert (oracle.mesh_to_sdf(V2, F, dims, origin, spacing, raw=True) == FAR).all()


def test_flood_exterior_and_cavity():
    """Far exterior samples +10, far samples deep inside a solid -10, a closed air cavity's +10,
    an empty mesh +10 everywhere. Catches: mutations cavity_sign and border_not_seed."""
    dims, origin, spacing = _unit(96)
    h = spacing[0]
    # a small box in a large grid: far exterior
    V, F = synth.box_mesh((0.45, 0.45, 0.45), (0.55, 0.55, 0.55))
    out = oracle.mesh_to_sdf(V, F, dims, origin, spacing)
    P = _grid_points(dims, origin, spacing)
    dt = _box_surface_dist(P, (0.45,) * 3, (0.55,) * 3)
    _, bd = _bins(dims, spacing)
    far = dt > 2.5 * float(min(bd)) + 6 * h
    assert far.sum() > 1000 and (out[far] == 10.0).all()
    # far samples deep inside a thick solid touch no positive sample: -10
    V, F = synth.box_mesh((0.05, 0.05, 0.05), (0.95, 0.95, 0.95))
    out = oracle.mesh_to_sdf(V, F, dims, origin, spacing)
    dt = _box_surface_dist(P, (0.05,) * 3, (0.95,) * 3)
    deep = np.all((P > 0.05) & (P < 0.95), -1) & (dt > 2.5 * float(min(bd)) + 6 * h)
    assert deep.sum() > 1000 and (out[deep] == -10.0).all()
    # a closed air cavity wider than the band: its band is positive (outside the solid), and
    # every positive sample seeds the flood (mesh_to_sdf_cuda.cu:376-383), so its far samples
    # are +10 like the outside's
    V, F = synth.hollow_box_mesh((0.05, 0.05, 0.05), (0.95, 0.95, 0.95), wall=0.1)
    out = oracle.mesh_to_sdf(V, F, dims, origin, spacing)
    dt = np.minimum(_box_surface_dist(P, (0.05,) * 3, (0.95,) * 3),
                    _box_surface_dist(P, (0.15,) * 3, (0.85,) * 3))
    in_cavity = np.all((P > 0.15) & (P < 0.85), -1)
    deep = in_cavity & (dt > 2.5 * float(min(bd)) + 6 * h)
    assert deep.sum() > 1000 and (out[deep] == 10.0).all()
    near = in_cavity & (dt < min(bd)) & (dt > 1e-4)
    assert (out[near] > 0).all()
    # no triangles: every sample far, all joined to the border
    out = oracle.mesh_to_sdf(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32), dims,
                             origin, spacing)
    assert (out == 10.0).all()


def test_relaxation_column_over_a_flat_face():
    """A column of samples over the centre of a large flat face: the 3-D relaxation reduces to a
    1-D Jacobi recurrence (lateral neighbours are equal), run here for 4 passes with the keep
    rule |f| <= 1.5 h. Catches: mutations relax_passes and relax_in_place."""
    n = 65
    dims, origin, spacing = _unit(n)
    h = np.float32(spacing[0])
    V, F = synth.box_mesh((0.05, 0.05, 0.05), (0.95, 0.95, 0.5))
    raw = oracle.mesh_to_sdf(V, F, dims, origin, spacing, raw=True)
    out = oracle.mesh_to_sdf(V, F, dims, origin, spacing)
    k0 = 32  # z = 0.5: the top face
    col = raw[:, 32, 32].copy()
    col[col == FAR] = np.where(np.arange(n)[col == FAR] > k0, 10.0, -10.0)
    col = col.astype(np.float32)
    keep = np.float32(h * np.float32(1.5))
    for _ in range(4):
        new = col.copy()
        for k in range(n):
            f = col[k]
            if abs(f) <= keep:
                continue
            m = np.float32(abs(f))
            for j in (k - 1, k + 1):
                if 0 <= j < n:
                    m = min(m, np.float32(np.float32(abs(col[j])) + h))
            m = min(m, np.float32(np.float32(abs(f)) + h))  # lateral neighbours (equal)
            new[k] = np.float32(-m if f < 0 else m)
        col = new
    assert np.array_equal(out[:, 32, 32].view(np.uint32), col.view(np.uint32))
    # each pass carries the band one sample farther: 4 samples past the raw band, then +10
    band = int((raw[k0 + 1:, 32, 32] != FAR).sum())
    top = out[k0 + 1:, 32, 32]
    assert band > 8 and int((top < 10.0).sum()) == band + 4 and (top[band + 4:] == 10.0).all()


@pytest.mark.parametrize("name", ["icosphere", "hollow_box"])
def test_voxelize_solid(name):
    """Solid voxels hold every voxel well inside the mesh, none well outside, and a closed cavity
    is filled; far samples take -10 in the solid and +10 outside. Catches: mutations
    voxel_reach and no_dilation."""
    if name == "icosphere":
        V, F = synth.icosphere(2, radius=0.3)
    else:
        V, F = synth.hollow_box_mesh((0.2, 0.2, 0.2), (0.8, 0.8, 0.8), wall=0.1)
    n, vres = 40, 48
    dims, origin, spacing = _unit(n)
    solid = oracle.m2s_voxels(V, F, dims, origin, spacing, vres).astype(bool)
    s = np.float32(np.float32(n) * np.float32(spacing[0])) / np.float32(vres)
    c = (np.arange(vres, dtype=np.float32) * s).astype(np.float64)
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    C = np.stack([X, Y, Z], -1)
    diag = float(s) * np.sqrt(3.0)
    if name == "icosphere":  # convex, within the sphere of radius 0.3 around its centre
        inside = _inside_convex(C, V, F)
        well_out = np.sqrt(((C - 0.5) ** 2).sum(-1)) > 0.3 + 3 * diag
    else:  # the filled box: the outer box's interior
        inside = np.all((C > 0.2) & (C < 0.8), -1)
        well_out = ~inside & (_box_surface_dist(C, (0.2,) * 3, (0.8,) * 3) > 3 * diag)
    assert solid[inside].all()
    assert not solid[well_out].any() and well_out.sum() > 1000
    out = oracle.mesh_to_sdf(V, F, dims, origin, spacing, voxelize_first=True, voxel_res=vres)
    raw = oracle.mesh_to_sdf(V, F, dims, origin, spacing, raw=True)
    far = raw == FAR
    assert np.array_equal(out[~far].view(np.uint32), raw[~far].view(np.uint32))
    P = _grid_points(dims, origin, spacing)
    vi = np.clip(np.floor((P.astype(np.float32) - 0) / s), 0, vres - 1).astype(int)
    sv = solid[vi[..., 2], vi[..., 1], vi[..., 0]]
    assert np.array_equal(out[far] == -10.0, sv[far]) and np.all(np.abs(out[far]) == 10.0)


def test_voxelize_growth_seals_a_small_hole():
    """A sphere with a hole two voxels across (its six triangles nearest the +x pole removed): the
    shell alone reaches 0.866 voxel diagonals from the triangles and leaves the hole open, the
    one-voxel growth closes it, so the flood stays outside and the interior is solid. Catches:
    mutation no_dilation."""
    V, F = synth.icosphere(3, radius=0.3)
    n, vres = 40, 48
    dims, origin, spacing = _unit(n)
    s = np.float32(np.float32(n) * np.float32(spacing[0])) / np.float32(vres)
    centroids = V[F].astype(np.float64).mean(1) - 0.5
    keep = np.linalg.norm(centroids - np.array([0.3, 0.0, 0.0]), axis=1) > 2.0 * float(s)
    assert (~keep).sum() == 6
    solid = oracle.m2s_voxels(V, F[keep], dims, origin, spacing, vres).astype(bool)
    c = (np.arange(vres, dtype=np.float32) * s).astype(np.float64)
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2)
    deep = r < 0.3 - 3 * float(s) * np.sqrt(3.0)
    assert deep.sum() > 1000 and solid[deep].all()
