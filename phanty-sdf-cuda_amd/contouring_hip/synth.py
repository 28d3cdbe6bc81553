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


# ---------------------------------------------------------------------------------------------
# Triangle meshes (inputs of mesh_to_sdf): vertices (n, 3) float32, triangles (m, 3) int32,
# counter-clockwise seen from outside (outward normals by the right-hand rule).
# ---------------------------------------------------------------------------------------------
def icosphere(level=3, radius=0.35, center=(0.5, 0.5, 0.5)):
    """Subdivided icosahedron (level 0: 20 triangles, each level x4), vertices projected to the
    sphere in FP64 and rounded to FP32."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    V = [np.array(v, np.float64) / np.linalg.norm(v) for v in V]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        mid = {}

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                m = V[a] + V[b]
                V.append(m / np.linalg.norm(m))
                mid[key] = len(V) - 1
            return mid[key]
        F2 = []
        for a, b, c in F:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            F2 += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = F2
    P = np.array(V) * radius + np.array(center, np.float64)
    return P.astype(np.float32), np.array(F, np.int32)


def box_mesh(lo=(0.3, 0.3, 0.3), hi=(0.7, 0.7, 0.7)):
    """Axis-aligned box as 12 triangles (two per face, of equal area)."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    V = np.array([(x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0),
                  (x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)], np.float32)
    quads = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (2, 3, 7, 6), (1, 2, 6, 5), (0, 4, 7, 3)]
    F = []
    for a, b, c, d in quads:
        F += [(a, b, c), (a, c, d)]
    return V, np.array(F, np.int32)


def torus_mesh(R=0.25, r=0.1, nu=48, nv=24, center=(0.5, 0.5, 0.5)):
    """Torus around the z axis (major radius R, minor r) on an nu x nv parameter grid."""
    u = np.arange(nu) * (2 * math.pi / nu)
    v = np.arange(nv) * (2 * math.pi / nv)
    U, W = np.meshgrid(u, v, indexing="ij")
    X = (R + r * np.cos(W)) * np.cos(U) + center[0]
    Y = (R + r * np.cos(W)) * np.sin(U) + center[1]
    Z = r * np.sin(W) + center[2]
    V = np.stack([X, Y, Z], -1).reshape(-1, 3).astype(np.float32)
    F = []
    for i in range(nu):
        for j in range(nv):
            a = i * nv + j
            b = ((i + 1) % nu) * nv + j
            c = ((i + 1) % nu) * nv + (j + 1) % nv
            d = i * nv + (j + 1) % nv
            F += [(a, b, c), (a, c, d)]
    return V, np.array(F, np.int32)


def hollow_box_mesh(lo=(0.2, 0.2, 0.2), hi=(0.8, 0.8, 0.8), wall=0.2):
    """A box with a closed inner cavity: the outer box and an inner box (inset by `wall`) with
    its triangles reversed (normals into the cavity)."""
    V1, F1 = box_mesh(lo, hi)
    V2, F2 = box_mesh(tuple(a + wall for a in lo), tuple(b - wall for b in hi))
    F2 = F2[:, ::-1] + len(V1)
    return np.concatenate([V1, V2]), np.ascontiguousarray(np.concatenate([F1, F2]), np.int32)


def random_soup(count=200, seed=99, size=0.08):
    """Seeded random triangles (not a closed surface) in [0.2, 0.8]^3 (SplitMix64, in order:
    centre xyz, then three corner offsets per triangle)."""
    rng = SplitMix64(seed)
    V = []
    for _ in range(count):
        c = [rng.uniform(0.2, 0.8) for _ in range(3)]
        for _ in range(3):
            V.append([c[d] + rng.uniform(-size, size) for d in range(3)])
    V = np.array(V, np.float32)
    return V, np.arange(3 * count, dtype=np.int32).reshape(-1, 3)


def quads_to_triangles(faces):
    """Quads (a, b, c, d) -> triangles (a, b, c), (a, c, d)."""
    q = np.asarray(faces, np.int32).reshape(-1, 4)
    t = np.empty((2 * len(q), 3), np.int32)
    t[0::2] = q[:, [0, 1, 2]]
    t[1::2] = q[:, [0, 2, 3]]
    return t
