"""Test grids: the reference's own test fixtures (src/cpp/dc_tests.cpp), the benchmark
configurations at small sizes, and edge cases (exact zeros, ragged dimensions, 64-bit word
boundaries). FP32 arithmetic follows the reference's host loops op by op."""
import numpy as np

from contouring_hip import synth
from contouring_hip.synth import Grid


def _axis(res, o=-1.0):
    v = np.float32(2.0) / np.float32(res - 1)
    return (np.float32(o) + np.arange(res, dtype=np.float32) * v).astype(np.float32), float(v)


def _cube(res, fn, name):
    a, v = _axis(res)
    z, y, x = np.meshgrid(a, a, a, indexing="ij")
    vals = fn(x, y, z).astype(np.float32).ravel()
    return Grid(vals, res, v, (-1.0, -1.0, -1.0), name)


def ref_sphere(res=32):
    """dc_tests.cpp:14-50 (also dc_cli --generate sphere)."""
    return _cube(res, lambda x, y, z: np.sqrt(x * x + y * y + z * z) - np.float32(0.5),
                 f"ref_sphere{res}")


def ref_box(res=32):
    """dc_tests.cpp:52-91."""
    def f(x, y, z):
        h = np.float32(0.5)
        return np.maximum(np.abs(x) - h, np.maximum(np.abs(y) - h, np.abs(z) - h))
    return _cube(res, f, f"ref_box{res}")


def ref_plane(res=16):
    """dc_tests.cpp:93-129: value z."""
    return _cube(res, lambda x, y, z: z.copy(), f"ref_plane{res}")


def ref_thin_gap(res=32):
    """dc_tests.cpp:131-169: two r=0.25 spheres at x=+-0.2."""
    def f(x, y, z):
        a, b, r = np.float32(0.2), np.float32(0.25), None
        d1 = np.sqrt((x - a) * (x - a) + y * y + z * z) - b
        d2 = np.sqrt((x + a) * (x + a) + y * y + z * z) - b
        return np.minimum(d1, d2)
    return _cube(res, f, f"ref_thin_gap{res}")


def ref_const(res, value):
    """dc_tests.cpp:171-215: empty (+1) and full (-1) grids."""
    return _cube(res, lambda x, y, z: np.full_like(x, np.float32(value)), f"ref_const{value}")


def zeros_grid(nx=9, ny=8, nz=7, seed=3):
    """Random {-1, 0, +1} * U(0.5, 1.5) samples: many exact zeros (fa*fb == 0 edges)."""
    rng = np.random.default_rng(seed)
    s = rng.integers(-1, 2, size=nx * ny * nz).astype(np.float32)
    m = rng.uniform(0.5, 1.5, size=nx * ny * nz).astype(np.float32)
    g = Grid((s * m).astype(np.float32), nx, 0.125, (0.0, 0.0, 0.0), f"zeros_{nx}x{ny}x{nz}")
    g.nx, g.ny, g.nz = nx, ny, nz
    return g


def ragged_sphere(nx, ny, nz, r=0.3, name=None):
    """Sphere on a non-cubic grid with spacing 1/(n-1) per axis and a non-zero origin."""
    ax = [np.float32(1.0) / np.float32(n - 1) for n in (nx, ny, nz)]
    o = np.float32(-0.05)
    xs = [(o + np.arange(n, dtype=np.float32) * a).astype(np.float32) for n, a in zip((nx, ny, nz), ax)]
    z, y, x = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
    c = np.float32(0.45)
    vals = (np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2) - np.float32(r)).astype(np.float32)
    g = Grid(vals.ravel(), nx, float(ax[0]), (float(o), float(o), float(o)),
             name or f"ragged_{nx}x{ny}x{nz}")
    g.nx, g.ny, g.nz = nx, ny, nz
    g.vx, g.vy, g.vz = float(ax[0]), float(ax[1]), float(ax[2])
    return g


def small_suite():
    """Grids small enough for the oracle in well under a second each."""
    return [
        ref_sphere(32), ref_box(32), ref_plane(16), ref_thin_gap(32), ref_const(16, 1.0),
        ref_const(16, -1.0), zeros_grid(), zeros_grid(5, 5, 5, seed=7), ref_sphere(2),
        ragged_sphere(13, 7, 9), ragged_sphere(66, 20, 17), ragged_sphere(130, 9, 70),
        synth.sphere_unit(64), synth.cli_generate("sphere", 64), synth.csg8(48, seed=11),
        synth.blend64(40, seed=5),
    ]


# ---------------------------------------------------------------------------------------------
# Grids for the sparse brick / multi-vertex backends: thin sheets (same-sign edges with opposed
# gradients), near-zero noise (many thin edges per cell, the 12-entry cap), a tilted plane, and
# non-cubic grids with unequal spacing.
# ---------------------------------------------------------------------------------------------
def _coords(nx, ny, nz, spacing, origin):
    xs = [(np.float32(o) + np.arange(n, dtype=np.float32) * np.float32(h)).astype(np.float32)
          for n, h, o in zip((nx, ny, nz), spacing, origin)]
    z, y, x = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
    return x, y, z


def _grid(vals, dims, spacing, origin, name):
    g = Grid(np.ascontiguousarray(vals, np.float32).ravel(), dims[0], float(spacing[0]),
             tuple(float(o) for o in origin), name)
    g.nx, g.ny, g.nz = dims
    g.vx, g.vy, g.vz = (float(np.float32(h)) for h in spacing)
    return g


def slab_grid(n=24, scale=2.0, half=0.3, ball=True):
    """A sheet thinner than a cell: scale * (|z - c| - half * h) with c in the middle of a cell
    layer, so both samples of every z edge through it are positive, within 0.45 h of zero, with
    central differences (0, 0, -scale) and (0, 0, +scale): thin edges. `ball` adds a small sphere
    away from the sheet (the brick test needs a sign change)."""
    h = np.float32(1.0) / np.float32(n - 1)
    x, y, z = _coords(n, n, n, (h, h, h), (0.0, 0.0, 0.0))
    k0 = n // 2
    c = np.float32(k0) * h + np.float32(0.5) * h
    f = np.float32(scale) * (np.abs(z - c) - np.float32(half) * h)
    if ball:
        s = np.sqrt((x - np.float32(0.2)) ** 2 + (y - np.float32(0.2)) ** 2 +
                    (z - np.float32(0.15)) ** 2) - np.float32(0.08)
        f = np.minimum(f, s)
    return _grid(f, (n, n, n), (h, h, h), (0.0, 0.0, 0.0), f"slab{n}")


def near_noise_grid(n=20, seed=5, neg=0.15):
    """Samples in (0, 0.45 h] with a fraction `neg` negated: most edges are same-sign and near
    zero with random gradients (thin edges, the entry cap), the rest cross."""
    rng = np.random.default_rng(seed)
    h = np.float32(1.0) / np.float32(n - 1)
    v = rng.uniform(0.02, 0.45, size=n ** 3).astype(np.float32) * h
    v[rng.random(n ** 3) < neg] *= np.float32(-1.0)
    return _grid(v, (n, n, n), (h, h, h), (0.0, 0.0, 0.0), f"near_noise{n}_{seed}")


def tilted_plane(n=20, normal=(0.3, -0.5, 0.8), d=0.1):
    """The linear field normal . x - d (normal not normalized)."""
    h = np.float32(1.0) / np.float32(n - 1)
    x, y, z = _coords(n, n, n, (h, h, h), (-0.5, -0.5, -0.5))
    a, b, c = (np.float32(t) for t in normal)
    f = a * x + b * y + c * z - np.float32(d)
    return _grid(f, (n, n, n), (h, h, h), (-0.5, -0.5, -0.5), f"tilted_plane{n}")


def crossed_sheets(n=28):
    """Two thin sheets crossing at right angles plus a ball: cells where thin edges of both sheets
    meet (many entries, split decisions on mixed normals)."""
    h = np.float32(1.0) / np.float32(n - 1)
    x, y, z = _coords(n, n, n, (h, h, h), (0.0, 0.0, 0.0))
    cz = np.float32(n // 2) * h + np.float32(0.5) * h
    cx = np.float32(n // 3) * h + np.float32(0.5) * h
    a = np.float32(2.0) * (np.abs(z - cz) - np.float32(0.3) * h)
    b = np.float32(2.0) * (np.abs(x - cx) - np.float32(0.3) * h)
    s = np.sqrt((x - np.float32(0.8)) ** 2 + (y - np.float32(0.8)) ** 2 +
                (z - np.float32(0.2)) ** 2) - np.float32(0.1)
    return _grid(np.minimum(np.minimum(a, b), s), (n, n, n), (h, h, h), (0.0, 0.0, 0.0),
                 f"crossed_sheets{n}")


def uneven_blob(nx=37, ny=50, nz=29, spacing=(0.031, 0.021, 0.029), origin=(-0.05, 0.02, 0.1)):
    """A smooth blob (sphere plus a low-frequency wobble) on a non-cubic grid with unequal spacing
    and an offset origin."""
    x, y, z = _coords(nx, ny, nz, spacing, origin)
    c = (np.float32(0.55), np.float32(0.55), np.float32(0.5))
    r = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)
    f = r - np.float32(0.3) - np.float32(0.04) * np.sin(np.float32(9.0) * x) * np.cos(
        np.float32(7.0) * y)
    return _grid(f.astype(np.float32), (nx, ny, nz), spacing, origin, f"uneven_blob{nx}x{ny}x{nz}")


def flat_zero_grid(n=16):
    """Exact zeros over the lower half (max(z - 0.5, 0)) and a negative ball in the upper half:
    cells whose corners are all zero have zero normals, so their QEF drops every singular value
    (the mass-point fallback)."""
    h = np.float32(1.0) / np.float32(n - 1)
    x, y, z = _coords(n, n, n, (h, h, h), (0.0, 0.0, 0.0))
    f = np.maximum(z - np.float32(0.5), np.float32(0.0))
    s = np.sqrt((x - np.float32(0.5)) ** 2 + (y - np.float32(0.5)) ** 2 +
                (z - np.float32(0.8)) ** 2) - np.float32(0.12)
    return _grid(np.minimum(f, s), (n, n, n), (h, h, h), (0.0, 0.0, 0.0), f"flat_zero{n}")


def sparse_suite():
    """Grids for the sparse backends' parity tests (the oracle takes well under a second each)."""
    return small_suite() + [
        slab_grid(24), slab_grid(17, scale=3.0, half=0.35), near_noise_grid(20, 5),
        near_noise_grid(13, 9, neg=0.4), tilted_plane(20), crossed_sheets(28), uneven_blob(),
        flat_zero_grid(16),
    ]
