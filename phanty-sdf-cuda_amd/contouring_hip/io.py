"""Grid and mesh files of the reference's tools.

.sdf (read by dc_cli, src/cpp/dc_cli.cpp:31-47; written by scripts/test_glb_benchmark.py:20-25):
int32 nx, ny, nz; float32 vx, vy, vz; float32 ox, oy, oz; then nx*ny*nz float32 values, x fastest.
OBJ (dc_cli.cpp:17-29, test_glb_benchmark.py:13-18): "v x y z" lines, then "f a b c d" quads with
1-based indices.
"""
import numpy as np

from .backend import DenseSdfGrid

_HEADER = 12 + 12 + 12


def save_sdf(path, grid):
    """Writes a DenseSdfGrid (or any object with the same fields) as a .sdf file."""
    vals = np.ascontiguousarray(np.asarray(grid.values, dtype=np.float32).reshape(-1))
    if vals.size != int(grid.nx) * int(grid.ny) * int(grid.nz):
        raise ValueError("values size does not match nx*ny*nz")
    with open(path, "wb") as f:
        f.write(np.array([grid.nx, grid.ny, grid.nz], dtype=np.int32).tobytes())
        f.write(np.array([grid.vx, grid.vy, grid.vz], dtype=np.float32).tobytes())
        f.write(np.array([grid.ox, grid.oy, grid.oz], dtype=np.float32).tobytes())
        f.write(vals.tobytes())


def load_sdf(path):
    """Reads a .sdf file into a DenseSdfGrid (values: a float32 NumPy array). Raises ValueError on
    a short file or dimensions below 2, where dc_cli refuses the file too."""
    with open(path, "rb") as f:
        head = f.read(_HEADER)
        if len(head) != _HEADER:
            raise ValueError(f"{path}: truncated .sdf header")
        nx, ny, nz = (int(x) for x in np.frombuffer(head[:12], np.int32))
        vx, vy, vz = (float(x) for x in np.frombuffer(head[12:24], np.float32))
        ox, oy, oz = (float(x) for x in np.frombuffer(head[24:36], np.float32))
        if nx < 2 or ny < 2 or nz < 2:
            raise ValueError(f"{path}: grid dimensions below 2")
        n = nx * ny * nz
        vals = np.frombuffer(f.read(4 * n), np.float32)
        if vals.size != n:
            raise ValueError(f"{path}: {vals.size} of {n} values")
    return DenseSdfGrid(nx, ny, nz, vx, vy, vz, ox, oy, oz, vals.copy())


def save_obj(path, vertices, faces):
    """Writes flat vertices (3 per vertex) and quads (4 per face, 0-based) as an OBJ file."""
    V = np.asarray(vertices, np.float32).reshape(-1, 3)
    F = np.asarray(faces, np.int64).reshape(-1, 4) + 1
    with open(path, "w") as f:
        for x, y, z in V.tolist():
            f.write(f"v {x} {y} {z}\n")
        for a, b, c, d in F.tolist():
            f.write(f"f {a} {b} {c} {d}\n")


def load_obj(path):
    """Reads "v" and "f" lines of an OBJ file: (V (n, 3) float32, F (m, k) int32, 0-based);
    every face must have the same number of indices (4 for the tools' quads)."""
    V, F = [], []
    with open(path) as f:
        for line in f:
            if line.startswith("v "):
                V.append([float(t) for t in line.split()[1:4]])
            elif line.startswith("f "):
                F.append([int(t.split("/")[0]) - 1 for t in line.split()[1:]])
    return (np.asarray(V, np.float32).reshape(-1, 3),
            np.asarray(F, np.int32).reshape(len(F), -1) if F else np.zeros((0, 4), np.int32))
