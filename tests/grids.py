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
