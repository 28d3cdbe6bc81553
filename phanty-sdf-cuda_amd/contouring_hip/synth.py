"""Seeded synthetic SDF grids on the unit cube (and the dc_cli [-1,1]^3 generators).

Every random draw comes from SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
generators", OOPSLA 2014) mapped to U[0,1) as (z >> 11) * 2^-53, in the order documented per
generator. Rotations are Haar-uniform via uniform unit quaternions (Shoemake, "Uniform random
rotations", Graphics Gems III, 1992). Evaluation runs in FP64 in libdc_synth and is rounded to
FP32. Primitive formulas follow the reference's src/cpp/sdf.cpp (box :46-78, sphere :101-116,
capped cylinder :119-155).
"""
import ctypes
import math

import numpy as np

from ._lib import load_synth

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK64

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        u = (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def choice(self, seq):
        return seq[min(int(self.uniform() * len(seq)), len(seq) - 1)]

    def rotation(self):
        """Shoemake: u1,u2,u3 ~ U[0,1) -> unit quaternion -> rotation matrix (row-major)."""
        u1, u2, u3 = self.uniform(), self.uniform(), self.uniform()
        a, b = math.sqrt(1.0 - u1), math.sqrt(u1)
        x, y = a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2)
        z, w = b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3)
        return [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]


IDENT = [1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0]
OP_MIN, OP_SUB, OP_SMIN = 0, 1, 2


def sphere(c, r):
    return [0.0, *c, *IDENT, r, 0.0, 0.0]


def box(c, R, half):
    return [1.0, *c, *R, *half]


def cylinder(c, R, r, he):
    return [2.0, *c, *R, r, he, 0.0]


class Grid:
    """A dense FP32 SDF grid with the DenseSdfGrid fields (dims, spacing, origin)."""

    def __init__(self, values, n, spacing, origin, name, params=None):
        self.values = values
        self.nx = self.ny = self.nz = n
        self.vx = self.vy = self.vz = spacing
        self.ox, self.oy, self.oz = origin
        self.name = name
        self.params = params or {}

    @property
    def h(self):
        return self.vx


def _eval(n, prims, ops, k=0.0, clampv=0.0):
    lib = load_synth()
    out = np.empty(n * n * n, dtype=np.float32)
    P = np.ascontiguousarray(np.array(prims, dtype=np.float64).reshape(-1, 16))
    O = np.ascontiguousarray(np.array(ops, dtype=np.int32))
    rc = lib.dc_synth_csg(n, len(prims), P.ctypes.data_as(ctypes.c_void_p),
                          O.ctypes.data_as(ctypes.c_void_p), ctypes.c_double(k),
                          ctypes.c_double(clampv), out.ctypes.data_as(ctypes.c_void_p))
    if rc != 0:
        raise ValueError("dc_synth_csg rejected the parameters")
    return out


def unit_grid(n, values, name, params=None):
    h = float(np.float32(1.0 / (n - 1)))
    return Grid(values, n, h, (0.0, 0.0, 0.0), name, params)


def sphere_unit(n=64):
    """Sphere r=0.35 centred at (0.5,0.5,0.5) on the unit cube."""
    prims = [sphere((0.5, 0.5, 0.5), 0.35)]
    return unit_grid(n, _eval(n, prims, [0]), f"sphere_unit_{n}", {"r": 0.35})


def cli_generate(kind, res):
    """dc_cli --generate {sphere,box,plane} -g res: [-1,1]^3, spacing 2/(res-1), FP32 loops."""
    kinds = {"sphere": 0, "box": 1, "plane": 2}
    lib = load_synth()
    out = np.empty(res * res * res, dtype=np.float32)
    if lib.dc_synth_cli(kinds[kind], res, out.ctypes.data_as(ctypes.c_void_p)) != 0:
        raise ValueError("bad resolution")
    v = float(np.float32(2.0) / np.float32(res - 1))
    return Grid(out, res, v, (-1.0, -1.0, -1.0), f"cli_{kind}_{res}")


def csg8_params(seed):
    """Eight rotated boxes (union) minus one sphere.

    Draw order: for each of 8 boxes: half extents (3 x U[0.08,0.25]), centre (3 x U[0.3,0.7]),
    rotation (3 uniforms); then the sphere radius U[0.1,0.2] and centre (3 x U[0.3,0.7]).
    """
    rng = SplitMix64(seed)
    prims, ops = [], []
    for b in range(8):
        half = [rng.uniform(0.08, 0.25) for _ in range(3)]
        c = [rng.uniform(0.3, 0.7) for _ in range(3)]
        R = rng.rotation()
        prims.append(box(c, R, half))
        ops.append(OP_MIN)
    r = rng.uniform(0.1, 0.2)
    c = [rng.uniform(0.3, 0.7) for _ in range(3)]
    prims.append(sphere(c, r))
    ops.append(OP_SUB)
    return prims, ops


def csg8(n, seed=1234):
    prims, ops = csg8_params(seed)
    return unit_grid(n, _eval(n, prims, ops), f"csg8_{n}_s{seed}", {"seed": seed})


def blend64_params(seed):
    """64 spheres, draw order per sphere: radius U[0.03,0.12], centre 3 x U[0.15,0.85]."""
    rng = SplitMix64(seed)
    prims = []
    for _ in range(64):
        r = rng.uniform(0.03, 0.12)
        c = [rng.uniform(0.15, 0.85) for _ in range(3)]
        prims.append(sphere(c, r))
    return prims, [OP_MIN] + [OP_SMIN] * 63


def blend64(n, seed=4321, k=0.02):
    """Smooth-min (polynomial, k=0.02, left fold in index order) of 64 spheres, clamped to +-4h."""
    prims, ops = blend64_params(seed)
    h = 1.0 / (n - 1)
    return unit_grid(n, _eval(n, prims, ops, k=k, clampv=4.0 * h), f"blend64_{n}_s{seed}",
                     {"seed": seed, "k": k})


BATCH_RES = (32, 48, 64, 96, 128, 192, 256)


def random_csg_params(rng):
    """1-6 primitives; per primitive: kind (uniform of sphere/box/cylinder), centre 3 x U[.25,.75],
    then sphere r U[0.1,0.3] | box half extents 3 x U[0.05,0.25] + rotation |
    cylinder r U[0.05,0.2], full height U[0.1,0.4] + rotation; op for primitives after the first:
    min with probability 1/2, else max(a,-b)."""
    count = 1 + min(int(rng.uniform() * 6), 5)
    prims, ops = [], []
    for q in range(count):
        kind = rng.choice(("sphere", "box", "cylinder"))
        c = [rng.uniform(0.25, 0.75) for _ in range(3)]
        if kind == "sphere":
            prims.append(sphere(c, rng.uniform(0.1, 0.3)))
        elif kind == "box":
            half = [rng.uniform(0.05, 0.25) for _ in range(3)]
            prims.append(box(c, rng.rotation(), half))
        else:
            r = rng.uniform(0.05, 0.2)
            he = rng.uniform(0.1, 0.4)
            prims.append(cylinder(c, rng.rotation(), r, he))
        if q == 0:
            ops.append(OP_MIN)
        else:
            ops.append(OP_MIN if rng.uniform() < 0.5 else OP_SUB)
    return prims, ops


def batch_specs(count=256, seed=777):
    """Resolutions (uniform over BATCH_RES) and CSG parameters for a batch of uneven grids."""
    rng = SplitMix64(seed)
    specs = []
    for _ in range(count):
        n = rng.choice(BATCH_RES)
        prims, ops = random_csg_params(rng)
        specs.append((n, prims, ops))
    return specs


def batch_grids(count=256, seed=777):
    return [unit_grid(n, _eval(n, prims, ops), f"batch{q}_{n}")
            for q, (n, prims, ops) in enumerate(batch_specs(count, seed))]
