"""ctypes access to the CPU restatement of the reference CpuDualContouringBackend.

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package. It checks the product; it is never the product.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB: a library built from a modified copy of dc_oracle.cpp (the mutation tests)
LIB = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "build", "liboracle_dc.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        raise RuntimeError(f"oracle not built: {LIB} (run make -C oracle)")
    lib = ctypes.CDLL(LIB)
    P, I, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
    lib.oracle_dc_run.restype = P
    lib.oracle_dc_run.argtypes = [I, I, I, F, F, F, F, F, F, P]
    lib.oracle_dc_run_f64.restype = P
    lib.oracle_dc_run_f64.argtypes = [I, I, I, F, F, F, F, F, F, P]
    lib.oracle_dc_free.argtypes = [P]
    lib.oracle_dc_counts.argtypes = [P, P, P, P]
    lib.oracle_dc_copy.argtypes = [P, P, P, P, P, P]
    lib.oracle_dc_hermite.argtypes = [P, I, P, P, P]
    lib.oracle_dc_backward.argtypes = [P, P, P]
    lib.oracle_sym3_eig.argtypes = [P, P, P]
    lib.oracle_qef_solve.argtypes = [I, P, P, P]
    lib.oracle_triangulate.argtypes = [P, ctypes.c_int64, P, ctypes.c_int64, P]
    lib.oracle_set_threads.argtypes = [I]
    i64 = ctypes.c_int64
    lib.oracle_postprocess.restype = P
    lib.oracle_postprocess.argtypes = [P, i64, P, i64, I, I]
    lib.oracle_post_counts.argtypes = [P, P, P, P]
    lib.oracle_post_copy.argtypes = [P, P, P]
    lib.oracle_post_free.argtypes = [P]
    lib.oracle_weld.restype = P
    lib.oracle_weld.argtypes = [P, i64, P, i64]
    m2s = [P, i64, P, i64, I, I, I, F, F, F, F, F, F]
    lib.oracle_mesh_to_sdf.restype = I
    lib.oracle_mesh_to_sdf.argtypes = m2s + [I, I, P]
    lib.oracle_mesh_to_sdf_raw.restype = I
    lib.oracle_mesh_to_sdf_raw.argtypes = m2s + [P]
    lib.oracle_m2s_voxels.restype = I
    lib.oracle_m2s_voxels.argtypes = m2s + [I, P]
    lib.oracle_m2s_sample_state.restype = None
    lib.oracle_m2s_sample_state.argtypes = m2s + [I, I, I, P]
    lib.oracle_point_triangle.restype = F
    lib.oracle_point_triangle.argtypes = [P, P, P, P, P]
    lib.oracle_sparse_run.restype = P
    lib.oracle_sparse_run.argtypes = [P, I, I, I, F, F, F, F, F, F, I, I, I]
    lib.oracle_sparse_stats.argtypes = [P, P]
    lib.oracle_sparse_copy.argtypes = [P, P, P, P, P, P]
    lib.oracle_sparse_free.argtypes = [P]
    lib.oracle_sparse_jacobi.restype = I
    lib.oracle_sparse_jacobi.argtypes = [P, P, F, P]
    lib.oracle_sparse_cell.restype = I
    lib.oracle_sparse_cell.argtypes = [P, I, I, I, F, F, F, F, F, F, I, I, I, I, P, P]
    _lib = lib
    return lib


def set_threads(n):
    """OpenMP threads of the oracle's forward loops (1 = the reference's single-threaded run)."""
    load().oracle_set_threads(int(n))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Result:
    """Outputs of one oracle extraction (the handle is kept for hermite data and the adjoint)."""

    def __init__(self, handle, dims):
        self._h = handle
        self.dims = dims
        lib = load()
        nv, nf, navg = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        lib.oracle_dc_counts(handle, ctypes.byref(nv), ctypes.byref(nf), ctypes.byref(navg))
        self.n_vertices, self.n_faces, self.averaged_edges = nv.value, nf.value, navg.value
        self.V32 = np.empty((self.n_vertices, 3), np.float32)
        self.V64 = np.empty((self.n_vertices, 3), np.float64)
        self.F = np.empty((self.n_faces, 4), np.int32)
        self.active = np.empty(self.n_vertices, np.int64)
        self.dots = np.empty(self.n_faces, np.float64)
        lib.oracle_dc_copy(handle, _p(self.V32), _p(self.V64), _p(self.F), _p(self.active),
                           _p(self.dots))

    def hermite(self, averaged=True):
        n = self.n_vertices
        mask = np.empty(n, np.uint16)
        pts = np.empty((n, 12, 3), np.float64)
        nrm = np.empty((n, 12, 3), np.float64)
        load().oracle_dc_hermite(self._h, 1 if averaged else 0, _p(mask), _p(pts), _p(nrm))
        return mask, pts, nrm

    def backward(self, gV):
        gV = np.ascontiguousarray(gV, dtype=np.float64).reshape(-1, 3)
        assert gV.shape[0] == self.n_vertices
        nx, ny, nz = self.dims
        out = np.empty(nx * ny * nz, np.float64)
        load().oracle_dc_backward(self._h, _p(gV), _p(out))
        return out

    def __del__(self):
        if getattr(self, "_h", None):
            load().oracle_dc_free(self._h)
            self._h = None


def run(values, nx, ny, nz, vx, vy, vz, ox=0.0, oy=0.0, oz=0.0, f64=False):
    lib = load()
    if f64:
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        h = lib.oracle_dc_run_f64(nx, ny, nz, vx, vy, vz, ox, oy, oz, _p(v))
    else:
        v = np.ascontiguousarray(values, dtype=np.float32).ravel()
        h = lib.oracle_dc_run(nx, ny, nz, vx, vy, vz, ox, oy, oz, _p(v))
    if not h:
        raise ValueError("oracle rejected the grid")
    return Result(h, (nx, ny, nz))


def run_grid(grid, f64=False):
    return run(grid.values, grid.nx, grid.ny, grid.nz, grid.vx, grid.vy, grid.vz, grid.ox,
               grid.oy, grid.oz, f64=f64)


def sym3_eig(M):
    M = np.ascontiguousarray(M, dtype=np.float64)
    lam = np.empty(3, np.float64)
    Q = np.empty((3, 3), np.float64)
    load().oracle_sym3_eig(_p(M), _p(lam), _p(Q))
    return lam, Q


def qef_solve(normals, points):
    n = np.ascontiguousarray(normals, dtype=np.float64)
    p = np.ascontiguousarray(points, dtype=np.float64)
    x = np.empty(3, np.float64)
    load().oracle_qef_solve(n.shape[0], _p(n), _p(p), _p(x))
    return x


def triangulate(F, V):
    """triangulate_v1 of the reference (contouring.cpp:88-113): quads (F x 4 int32) and vertices
    (A x 3 doubles) -> 2F x 3 int32 triangles, each oriented with its quad's normal."""
    F = np.ascontiguousarray(F, dtype=np.int32).reshape(-1, 4)
    V = np.ascontiguousarray(V, dtype=np.float64).reshape(-1, 3)
    T = np.empty((2 * F.shape[0], 3), np.int32)
    load().oracle_triangulate(_p(F), F.shape[0], _p(V), V.shape[0], _p(T))
    return T


POST_INFO = ("welded_vertices", "welded_faces", "components", "kept_faces", "boundary_edges",
             "loops_filled", "added_vertices", "added_faces")


def postprocess(V, F, close_holes=True, remove_floaters=True):
    """The reference's postprocess_mesh (src/cpp/dc_backend.cpp:217-558) restated on the CPU.

    V: (n, 3) float32, F: (m, 4) int32. Returns (V2, F2, info) with info a dict of POST_INFO.
    """
    lib = load()
    V = np.ascontiguousarray(V, np.float32).reshape(-1, 3)
    F = np.ascontiguousarray(F, np.int32).reshape(-1, 4)
    h = lib.oracle_postprocess(_p(V), V.shape[0], _p(F), F.shape[0], int(bool(close_holes)),
                               int(bool(remove_floaters)))
    if not h:
        raise RuntimeError("oracle postprocess: a hole fill did not converge")
    return _post_result(h)


def weld(V, F):
    """The reference's merge_vertices_in_place (src/cpp/dc_backend.cpp:127-215): (V2, F2)."""
    lib = load()
    V = np.ascontiguousarray(V, np.float32).reshape(-1, 3)
    F = np.ascontiguousarray(F, np.int32).reshape(-1, 4)
    V2, F2, _ = _post_result(lib.oracle_weld(_p(V), V.shape[0], _p(F), F.shape[0]))
    return V2, F2


def _post_result(h):
    lib = load()
    try:
        nv, nf = ctypes.c_int64(), ctypes.c_int64()
        info = np.zeros(8, np.int64)
        lib.oracle_post_counts(h, ctypes.byref(nv), ctypes.byref(nf), _p(info))
        V2 = np.empty((nv.value, 3), np.float32)
        F2 = np.empty((nf.value, 4), np.int32)
        lib.oracle_post_copy(h, _p(V2), _p(F2))
    finally:
        lib.oracle_post_free(h)
    return V2, F2, dict(zip(POST_INFO, (int(x) for x in info)))


def _mesh_args(V, F):
    V = np.ascontiguousarray(V, np.float32).reshape(-1, 3)
    F = np.ascontiguousarray(F, np.int32).reshape(-1, 3)
    return V, F


def mesh_to_sdf(V, F, dims, origin, spacing, voxelize_first=False, voxel_res=256, raw=False):
    """The reference's compute_mesh_sdf_device_cuda (src/cuda/mesh_to_sdf_cuda.cu:740-973)
    restated on the CPU: the sampled SDF of triangle mesh (V, F) as a (nz, ny, nx) float32 array.
    raw=True: the evaluation alone (FLT_MAX beyond 2.5 bins), before the flood and relaxation."""
    lib = load()
    V, F = _mesh_args(V, F)
    nx, ny, nz = (int(d) for d in dims)
    out = np.empty((nz, ny, nx), np.float32)
    args = [_p(V), V.shape[0], _p(F), F.shape[0], nx, ny, nz, *(float(o) for o in origin),
            *(float(v) for v in spacing)]
    if raw:
        rc = lib.oracle_mesh_to_sdf_raw(*args, _p(out))
    else:
        rc = lib.oracle_mesh_to_sdf(*args, int(bool(voxelize_first)), int(voxel_res), _p(out))
    if rc != 0:
        raise ValueError("oracle mesh_to_sdf: a triangle names a vertex out of range")
    return out


def m2s_voxels(V, F, dims, origin, spacing, voxel_res=256):
    """The voxelize option's solid voxels as a (vres, vres, vres) uint8 array (z, y, x)."""
    lib = load()
    V, F = _mesh_args(V, F)
    nx, ny, nz = (int(d) for d in dims)
    args = [_p(V), V.shape[0], _p(F), F.shape[0], nx, ny, nz, *(float(o) for o in origin),
            *(float(v) for v in spacing), int(voxel_res)]
    vres = lib.oracle_m2s_voxels(*args, None)
    solid = np.empty((vres, vres, vres), np.uint8)
    lib.oracle_m2s_voxels(*args, _p(solid))
    return solid


def point_triangle(p, a, b, c):
    """(squared distance, closest point) of p to triangle (a, b, c), the oracle's FP32 routine."""
    lib = load()
    arrs = [np.ascontiguousarray(x, np.float32).reshape(3) for x in (p, a, b, c)]
    q = np.empty(3, np.float32)
    d = lib.oracle_point_triangle(*(_p(x) for x in arrs), _p(q))
    return float(d), q


def m2s_sample_state(V, F, dims, origin, spacing, ix, iy, iz):
    """One sample's final (squared distance, closest point, summed weighted normal)."""
    lib = load()
    V, F = _mesh_args(V, F)
    out = np.empty(7, np.float32)
    lib.oracle_m2s_sample_state(_p(V), V.shape[0], _p(F), F.shape[0], *(int(d) for d in dims),
                                *(float(o) for o in origin), *(float(v) for v in spacing),
                                int(ix), int(iy), int(iz), _p(out))
    return float(out[0]), out[1:4].copy(), out[4:7].copy()


# stats of the sparse brick / multi-vertex extraction, in the reference's DualContouringStats names
SPARSE_STATS = ("total_cells", "active_cells", "total_bricks", "active_bricks", "vertex_count",
                "face_count", "qef_fallback_count", "clamp_count", "invalid_count",
                "ambiguous_cells", "multi_vertex_cells", "one_vertex_cells", "two_vertex_cells",
                "split_rejection_count", "bad_qef_count", "faces_skipped_due_to_missing_cluster")


def sparse(values, nx, ny, nz, vx, vy, vz, ox=0.0, oy=0.0, oz=0.0, brick_size=8, normal_mode=0,
           multi_vertex=False):
    """The reference's CudaSparse[Mvdc]DualContouringBackend on a grid (values x fastest).

    Returns a dict: vertices (n, 3) float32, vnormals (n, 3) (split cells' vertex normals, else
    zero), faces (m, 4) int32, active (linear cell ids, ascending), vcount (vertices per active
    cell) and stats (SPARSE_STATS)."""
    lib = load()
    vals = np.ascontiguousarray(np.asarray(values, np.float32).reshape(-1))
    if vals.size != nx * ny * nz:
        raise ValueError("values do not match the grid size")
    h = lib.oracle_sparse_run(_p(vals), nx, ny, nz, vx, vy, vz, ox, oy, oz, int(brick_size),
                              int(normal_mode), int(bool(multi_vertex)))
    if not h:
        raise ValueError("invalid sparse extraction arguments")
    try:
        st = np.zeros(16, np.int64)
        lib.oracle_sparse_stats(h, _p(st))
        nv, nf, na = int(st[4]), int(st[5]), int(st[1])
        V = np.empty((nv, 3), np.float32)
        N = np.empty((nv, 3), np.float32)
        Fc = np.empty((nf, 4), np.int32)
        A = np.empty(na, np.int64)
        C = np.empty(na, np.int32)
        lib.oracle_sparse_copy(h, _p(V), _p(N), _p(Fc), _p(A), _p(C))
    finally:
        lib.oracle_sparse_free(h)
    return {"vertices": V, "vnormals": N, "faces": Fc, "active": A, "vcount": C,
            "stats": dict(zip(SPARSE_STATS, (int(x) for x in st)))}


def sparse_grid(grid, **kw):
    return sparse(grid.values, grid.nx, grid.ny, grid.nz, grid.vx, grid.vy, grid.vz, grid.ox,
                  grid.oy, grid.oz, **kw)


def sparse_jacobi(B, d, thr=0.01):
    """The sparse backends' 3x3 Jacobi pseudo-inverse solve: (y, singular)."""
    B = np.ascontiguousarray(np.asarray(B, np.float32).reshape(9))
    d = np.ascontiguousarray(np.asarray(d, np.float32).reshape(3))
    y = np.zeros(3, np.float32)
    s = load().oracle_sparse_jacobi(_p(B), _p(d), float(thr), _p(y))
    return y, bool(s)


def sparse_cell(values, dims, spacing, origin, ijk, normal_mode=0):
    """One cell's hermite entries (points, normals), (m, 3) float32 each."""
    vals = np.ascontiguousarray(np.asarray(values, np.float32).reshape(-1))
    P = np.zeros((12, 3), np.float32)
    Nn = np.zeros((12, 3), np.float32)
    m = load().oracle_sparse_cell(_p(vals), *(int(x) for x in dims), *spacing, *origin,
                                  *(int(x) for x in ijk), int(normal_mode), _p(P), _p(Nn))
    return P[:m], Nn[:m]
