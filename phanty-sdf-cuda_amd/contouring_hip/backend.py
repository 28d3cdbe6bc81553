"""Python mirror of the reference backend classes, on top of the C ABI (include/dc_hip.h).

Reference interface (file:line in the reference tree):
  DenseSdfGrid, DualContouringMesh, DualContouringStats   include/dc_backend.h:13-78,
                                                          src/cpp/bindings.cpp:135-185
  <Backend>().extract(grid) -> (mesh, stats)              src/cpp/bindings.cpp:186-222
  DenseSdfGridDevice, extract_device(grid)                include/dc_cuda.h:24-35,
                                                          src/cpp/bindings.cpp:204-227
  mesh_to_sdf_cuda / mesh_to_sdf_device_cuda(vertices, faces, nx, ny, nz, ox, oy, oz, vx, vy,
  vz) -> (grid, ms)                                       src/cpp/bindings.cpp:270-284
  NormalComputationMode                                   include/dc_cuda.h:5-9,
                                                          src/cpp/bindings.cpp:209-213
  CudaSparse[Mvdc]DualContouringBackend(brick_size, normal_mode, chunk_size[,
  multi_vertex_cells]).extract / extract_device           include/dc_cuda.h:37-62,
                                                          src/cpp/bindings.cpp:229-262
HipDualContouringBackend takes the place of CudaDualContouringBackend,
HipSparseDualContouringBackend and HipSparseMvdcDualContouringBackend the places of the sparse
pair. There is no CPU fallback: without the HIP library or a device every call raises.
"""
import ctypes

import numpy as np

from . import _lib

PRECISIONS = {"plain-fp64": _lib.DC_HIP_PLAIN_FP64, "fp32-fallback": _lib.DC_HIP_FP32_FALLBACK}


class DenseSdfGrid:
    """include/dc_backend.h:13-21. values: nx*ny*nz floats, index i + nx*(j + ny*k)."""

    def __init__(self, nx=0, ny=0, nz=0, vx=0.0, vy=0.0, vz=0.0, ox=0.0, oy=0.0, oz=0.0,
                 values=None):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.vx, self.vy, self.vz = vx, vy, vz
        self.ox, self.oy, self.oz = ox, oy, oz
        self.values = values if values is not None else []


class DenseSdfGridDevice:
    """include/dc_cuda.h:24-29: d_values is a device pointer (int) or a float32 torch tensor."""

    def __init__(self, nx=0, ny=0, nz=0, vx=0.0, vy=0.0, vz=0.0, ox=0.0, oy=0.0, oz=0.0,
                 d_values=None):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.vx, self.vy, self.vz = vx, vy, vz
        self.ox, self.oy, self.oz = ox, oy, oz
        self.d_values = d_values


class DualContouringMesh:
    """include/dc_backend.h:23-26: vertices (3 floats per vertex), faces (4 ints per quad)."""

    def __init__(self, vertices=None, faces=None, vertices_f64=None):
        self.vertices = vertices if vertices is not None else np.zeros(0, np.float32)
        self.faces = faces if faces is not None else np.zeros(0, np.int32)
        self.vertices_f64 = vertices_f64


_STATS_FIELDS = (
    "backend", "nx", "ny", "nz", "total_cells", "active_cells", "total_bricks", "active_bricks",
    "upload_ms", "marking_ms", "compaction_ms", "active_brick_marking_ms",
    "active_brick_compaction_ms", "active_cell_marking_ms", "active_cell_compaction_ms", "qef_ms",
    "face_emission_ms", "face_count_ms", "face_prefix_sum_ms", "face_fill_ms", "download_ms",
    "total_ms", "vertex_count", "face_count", "qef_fallback_count", "clamp_count", "invalid_count",
    "ambiguous_cells", "multi_vertex_cells", "one_vertex_cells", "two_vertex_cells",
    "split_rejection_count", "bad_qef_count", "faces_skipped_due_to_missing_cluster",
)


class DualContouringStats:
    """include/dc_backend.h:44-78 (every field, zero-initialised) plus backward_ms."""

    def __init__(self):
        for f in _STATS_FIELDS:
            setattr(self, f, "" if f == "backend" else 0)
        self.backward_ms = 0.0
        self.normals_ms = 0.0

    @classmethod
    def from_c(cls, s, backend="hip"):
        st = cls()
        st.backend = backend
        for f in ("nx", "ny", "nz", "total_cells", "active_cells", "upload_ms", "marking_ms",
                  "compaction_ms", "face_prefix_sum_ms", "qef_ms", "face_fill_ms",
                  "face_emission_ms", "download_ms", "total_ms", "backward_ms", "vertex_count",
                  "face_count", "qef_fallback_count", "invalid_count"):
            setattr(st, f, getattr(s, f))
        st.active_cell_marking_ms = s.marking_ms
        st.active_cell_compaction_ms = s.compaction_ms
        st.normals_ms = s.normals_ms
        st.graph_mode = s.graph_mode
        return st

    def as_dict(self):
        return {f: getattr(self, f) for f in _STATS_FIELDS + ("backward_ms", "normals_ms")}


def _ptr(x):
    """Device pointer of a torch tensor or an int."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def _check_tensor(t, name, numel, device, dtype="float32"):
    """Raises ValueError unless t is a contiguous CUDA tensor of `dtype` with `numel` elements on
    device index `device`: its data pointer is handed to kernels that index it by those sizes."""
    import torch
    want = getattr(torch, dtype)
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor (or an int device pointer)")
    if t.dtype != want:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_cuda or (t.device.index or 0) != device:
        raise ValueError(f"{name} must live on cuda:{device}, got {t.device}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if t.numel() != numel:
        raise ValueError(f"{name} has {t.numel()} elements, expected {numel}")


def _mesh_array(x, name, per, dtype, device):
    """(pointer, item count, the array kept alive) of a mesh array: a CUDA tensor on `device`, or
    anything NumPy converts (CPU tensors included); `per` values per item."""
    if hasattr(x, "data_ptr") and getattr(x, "is_cuda", False):
        import torch
        t = x.contiguous()
        if t.dtype != getattr(torch, dtype):
            raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
        if (t.device.index or 0) != device:
            raise ValueError(f"{name} must live on cuda:{device}, got {t.device}")
        if t.numel() % per:
            raise ValueError(f"{name} must hold {per} values per item")
        return ctypes.c_void_p(t.data_ptr()), t.numel() // per, t
    if hasattr(x, "detach"):
        x = x.detach().cpu().numpy()
    a = np.ascontiguousarray(np.asarray(x, dtype).reshape(-1))
    if a.size % per:
        raise ValueError(f"{name} must hold {per} values per item")
    return a.ctypes.data_as(ctypes.c_void_p), a.size // per, a


class HipDualContouringBackend:
    """MI355X dense dual contouring (replaces CudaDualContouringBackend, dc_cuda.h:31-35).

    precision: "plain-fp64" (the CPU backend's arithmetic, FP64 vertices available) or
    "fp32-fallback" (FP32 assembly, FP64 only for ill-conditioned cells).
    """

    def __init__(self, precision="plain-fp64", device=0, mark_rows=None, mark_waves=None):
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        self.precision = precision
        self.device = device
        self._lib = _lib.load_hip()
        ctx = ctypes.c_void_p()
        rc = self._lib.dc_hip_create(device, ctypes.byref(ctx))
        if rc != _lib.DC_HIP_OK:
            raise _lib.DcHipError(rc, self._lib.dc_hip_status_string(rc).decode())
        self._ctx = ctx
        self.generation = 0
        self._samples = 0  # samples of the last forward (single grid, or the whole batch)
        # the device tensors the context's last forward points at: the adjoint reads them again
        self._inputs = ()
        if mark_rows is not None:
            self.set_option("mark_rows", mark_rows)
        if mark_waves is not None:
            self.set_option("mark_waves", mark_waves)

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx is not None and ctx.value:
            self._lib.dc_hip_destroy(ctx)
            self._ctx = None

    # -- plumbing ---------------------------------------------------------------------------
    def _check(self, rc):
        if rc != _lib.DC_HIP_OK:
            msg = self._lib.dc_hip_last_error(self._ctx).decode()
            raise _lib.DcHipError(rc, msg or self._lib.dc_hip_status_string(rc).decode())

    def set_option(self, key, value):
        self._check(self._lib.dc_hip_set_option(self._ctx, key.encode(), int(value)))
        self.__dict__.setdefault("_options", {})[key] = int(value)

    def option(self, key, default=None):
        """The value last set for `key` through set_option (else `default`)."""
        return self.__dict__.get("_options", {}).get(key, default)

    def autotune(self, grids, **kw):
        """Time every mark tile shape on `grids` (device grids) and keep the fastest; see
        contouring_hip.autotune.autotune_mark."""
        from .autotune import autotune_mark
        return autotune_mark(self, grids, **kw)

    def autotune_launch(self, grid, **kw):
        """Time the per-cell kernels' launch shapes on `grid` (a device grid) by whole steps and
        keep the fastest; see contouring_hip.autotune.autotune_launch."""
        from .autotune import autotune_launch
        return autotune_launch(self, grid, **kw)

    def set_stream(self, stream_handle):
        self._check(self._lib.dc_hip_set_stream(self._ctx, ctypes.c_void_p(stream_handle or 0)))

    def _use_torch_stream(self):
        import torch
        self.set_stream(torch.cuda.current_stream(self.device).cuda_stream)

    @staticmethod
    def _cgrid(g, ptr):
        return _lib.DcHipGrid(int(g.nx), int(g.ny), int(g.nz), float(g.vx), float(g.vy),
                              float(g.vz), float(g.ox), float(g.oy), float(g.oz),
                              ctypes.c_void_p(ptr))

    def counts(self):
        nv, nf = ctypes.c_int64(), ctypes.c_int64()
        self._check(self._lib.dc_hip_result_counts(self._ctx, ctypes.byref(nv), ctypes.byref(nf)))
        return nv.value, nf.value

    def _download(self, with_f64):
        nv, nf = self.counts()
        V = np.empty(3 * nv, np.float32)
        F = np.empty(4 * nf, np.int32)
        V64 = np.empty(3 * nv, np.float64) if with_f64 else None
        self._check(self._lib.dc_hip_copy_result(
            self._ctx, V.ctypes.data_as(ctypes.c_void_p),
            V64.ctypes.data_as(ctypes.c_void_p) if with_f64 else None,
            F.ctypes.data_as(ctypes.c_void_p)))
        return DualContouringMesh(V, F, V64)

    # -- reference API ------------------------------------------------------------------------
    def extract(self, grid, with_f64=None):
        """Host grid -> (DualContouringMesh, DualContouringStats) (dc_cuda.cu:614-649)."""
        import time
        vals = np.ascontiguousarray(np.asarray(grid.values, dtype=np.float32))
        if vals.size != int(grid.nx) * int(grid.ny) * int(grid.nz):
            raise ValueError("values size does not match nx*ny*nz")
        cg = self._cgrid(grid, vals.ctypes.data)
        st = _lib.DcHipStats()
        self.set_stream(None)
        self._check(self._lib.dc_hip_extract(self._ctx, ctypes.byref(cg),
                                             PRECISIONS[self.precision], ctypes.byref(st)))
        self.generation += 1
        self._samples = vals.size
        self._inputs = ()  # the samples were uploaded into the context's own buffer
        t0 = time.perf_counter()
        if with_f64 is None:
            with_f64 = self.precision == "plain-fp64"
        mesh = self._download(with_f64)
        stats = DualContouringStats.from_c(st)
        stats.download_ms = (time.perf_counter() - t0) * 1e3
        stats.total_ms += stats.download_ms
        return mesh, stats

    def extract_device(self, grid, with_f64=None):
        """Device grid -> host mesh (dc_cuda.cu:651-847). grid: DenseSdfGridDevice whose
        d_values is a float32 CUDA tensor (checked: dtype, device, contiguity, nx*ny*nz
        elements) or an int device pointer (the caller vouches for nx*ny*nz floats)."""
        import time
        n = int(grid.nx) * int(grid.ny) * int(grid.nz)
        if not isinstance(grid.d_values, int):
            _check_tensor(grid.d_values, "d_values", n, self.device)
        st = _lib.DcHipStats()
        cg = self._cgrid(grid, _ptr(grid.d_values))
        if not isinstance(grid.d_values, int):
            self._use_torch_stream()
        self._check(self._lib.dc_hip_extract_device(self._ctx, ctypes.byref(cg),
                                                    PRECISIONS[self.precision], ctypes.byref(st)))
        self.generation += 1
        self._samples = n
        self._inputs = () if isinstance(grid.d_values, int) else (grid.d_values,)
        t0 = time.perf_counter()
        if with_f64 is None:
            with_f64 = self.precision == "plain-fp64"
        mesh = self._download(with_f64)
        stats = DualContouringStats.from_c(st)
        stats.download_ms = (time.perf_counter() - t0) * 1e3
        stats.total_ms += stats.download_ms
        return mesh, stats

    # -- tensor API (device in, device out) ---------------------------------------------------
    def extract_tensor(self, phi, spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0),
                       with_f64=False):
        """phi: contiguous float32 CUDA tensor of shape (nz, ny, nx). Returns (V [A,3] float32,
        F [F,4] int32[, V64 [A,3] float64]) as new CUDA tensors, plus the stats."""
        import torch
        if phi.dtype != torch.float32 or not phi.is_cuda or phi.dim() != 3:
            raise ValueError("phi must be a float32 CUDA tensor of shape (nz, ny, nx)")
        if (phi.device.index or 0) != self.device:
            raise ValueError(f"phi must live on cuda:{self.device}, got {phi.device}")
        phi = phi.contiguous()
        nz, ny, nx = phi.shape
        g = DenseSdfGridDevice(nx, ny, nz, *spacing, *origin, d_values=phi)
        self._use_torch_stream()
        st = _lib.DcHipStats()
        cg = self._cgrid(g, phi.data_ptr())
        self._check(self._lib.dc_hip_extract_device(self._ctx, ctypes.byref(cg),
                                                    PRECISIONS[self.precision], ctypes.byref(st)))
        self.generation += 1
        self._samples = phi.numel()
        self._inputs = (phi,)
        nv, nf = self.counts()
        V = torch.empty((nv, 3), dtype=torch.float32, device=phi.device)
        F = torch.empty((nf, 4), dtype=torch.int32, device=phi.device)
        V64 = torch.empty((nv, 3), dtype=torch.float64, device=phi.device) if with_f64 else None
        self._check(self._lib.dc_hip_copy_result(
            self._ctx, ctypes.c_void_p(V.data_ptr() if nv else 0),
            ctypes.c_void_p(V64.data_ptr() if (with_f64 and nv) else 0),
            ctypes.c_void_p(F.data_ptr() if nf else 0)))
        out = (V, F, V64) if with_f64 else (V, F)
        return out, DualContouringStats.from_c(st)

    def extract_batch(self, grids, with_f64=False):
        """List of DenseSdfGridDevice (tensor-backed) -> list of (V, F) CUDA tensor pairs.
        Every tensor is checked like extract_device's; a gradient for backward() after a batch
        is the concatenation of all grids' samples, grid-major."""
        import torch
        n = len(grids)
        if n == 0:
            raise ValueError("empty batch")
        total = 0
        for q, g in enumerate(grids):
            ng = int(g.nx) * int(g.ny) * int(g.nz)
            if not isinstance(g.d_values, int):
                _check_tensor(g.d_values, f"grids[{q}].d_values", ng, self.device)
            total += ng
        arr = (_lib.DcHipGrid * n)(*[self._cgrid(g, _ptr(g.d_values)) for g in grids])
        voff = np.zeros(n + 1, np.int64)
        foff = np.zeros(n + 1, np.int64)
        st = _lib.DcHipStats()
        self._use_torch_stream()
        self._check(self._lib.dc_hip_extract_batch(
            self._ctx, arr, n, PRECISIONS[self.precision], voff.ctypes.data_as(ctypes.c_void_p),
            foff.ctypes.data_as(ctypes.c_void_p), ctypes.byref(st)))
        self.generation += 1
        self._samples = total
        self._inputs = tuple(g.d_values for g in grids if not isinstance(g.d_values, int))
        nv, nf = self.counts()
        dev = torch.device("cuda", self.device)
        V = torch.empty((nv, 3), dtype=torch.float32, device=dev)
        F = torch.empty((nf, 4), dtype=torch.int32, device=dev)
        V64 = torch.empty((nv, 3), dtype=torch.float64, device=dev) if with_f64 else None
        self._check(self._lib.dc_hip_copy_result(
            self._ctx, ctypes.c_void_p(V.data_ptr() if nv else 0),
            ctypes.c_void_p(V64.data_ptr() if (with_f64 and nv) else 0),
            ctypes.c_void_p(F.data_ptr() if nf else 0)))
        out = []
        for q in range(n):
            a, b = int(voff[q]), int(voff[q + 1])
            c, d = int(foff[q]), int(foff[q + 1])
            item = (V[a:b], F[c:d]) + ((V64[a:b],) if with_f64 else ())
            out.append(item)
        return out, DualContouringStats.from_c(st)

    def triangles(self):
        """The last extraction's quads as 2F x 3 int32 triangles on the device, split and oriented
        like the reference's triangulate_v1 (src/cpp/contouring.cpp:88-113)."""
        import torch
        _, nf = self.counts()
        T = torch.empty((2 * nf, 3), dtype=torch.int32, device=torch.device("cuda", self.device))
        if nf:
            self._use_torch_stream()
            self._check(self._lib.dc_hip_triangulate(self._ctx, ctypes.c_void_p(T.data_ptr())))
        return T

    def postprocess(self, mesh=None, close_holes=True, remove_floaters=True, device=False):
        """postprocess_mesh (src/cpp/dc_backend.cpp:217-558) on the GPU: weld, keep the largest
        component (remove_floaters), close boundary loops (close_holes); the result equals the
        reference's. mesh: a DualContouringMesh whose vertices / faces are NumPy arrays (or
        sequences) or CUDA tensors, or None for this backend's last extraction (single grid), which
        stays valid for backward. Returns (DualContouringMesh, stats dict); device=True returns
        the arrays as CUDA tensors instead of NumPy."""
        st = _lib.DcHipPostStats()
        nv, nf = ctypes.c_int64(), ctypes.c_int64()
        if mesh is None:
            args = (None, 0, None, 0)
        else:
            vp, n_v, vkeep = _mesh_array(mesh.vertices, "vertices", 3, "float32", self.device)
            fp, n_f, fkeep = _mesh_array(mesh.faces, "faces", 4, "int32", self.device)
            args = (vp, n_v, fp, n_f)
            if not isinstance(vkeep, np.ndarray) or not isinstance(fkeep, np.ndarray):
                self._use_torch_stream()  # device inputs: ordered after the caller's work
        self._check(self._lib.dc_hip_postprocess(self._ctx, *args, int(bool(close_holes)),
                                                 int(bool(remove_floaters)), ctypes.byref(nv),
                                                 ctypes.byref(nf), ctypes.byref(st)))
        stats = {f: getattr(st, f) for f, _ in _lib.DcHipPostStats._fields_}
        if device:
            import torch
            dev = torch.device("cuda", self.device)
            Vo = torch.empty(3 * nv.value, dtype=torch.float32, device=dev)
            Fo = torch.empty(4 * nf.value, dtype=torch.int32, device=dev)
            self._use_torch_stream()
            self._check(self._lib.dc_hip_postprocess_copy(
                self._ctx, ctypes.c_void_p(Vo.data_ptr()), ctypes.c_void_p(Fo.data_ptr())))
        else:
            Vo = np.empty(3 * nv.value, np.float32)
            Fo = np.empty(4 * nf.value, np.int32)
            self._check(self._lib.dc_hip_postprocess_copy(
                self._ctx, Vo.ctypes.data_as(ctypes.c_void_p), Fo.ctypes.data_as(ctypes.c_void_p)))
        return DualContouringMesh(Vo, Fo), stats

    def mesh_to_sdf(self, vertices, faces, nx, ny, nz, ox, oy, oz, vx, vy, vz,
                    voxelize_first=False, voxel_res=256, out=None):
        """compute_mesh_sdf_cuda (src/cuda/mesh_to_sdf_cuda.cu:740-1002) on this backend's GPU:
        the signed distance of the triangle mesh sampled at (ox + i*vx, oy + j*vy, oz + k*vz).
        vertices: (n, 3) float32 and faces: (m, 3) int32, NumPy arrays (or sequences) or CUDA
        tensors. out: None for a new NumPy (nz, ny, nx) array, "device" for a new CUDA tensor of
        that shape, or a contiguous float32 CUDA tensor of nx*ny*nz elements written in place.
        Returns (values, stats dict); equal to the reference's result for the same input (each
        bin's triangles in ascending order, floods to their fixpoint)."""
        vp, n_v, vkeep = _mesh_array(vertices, "vertices", 3, "float32", self.device)
        fp, n_f, fkeep = _mesh_array(faces, "faces", 3, "int32", self.device)
        n = int(nx) * int(ny) * int(nz)
        if out is None:
            vals = np.empty((int(nz), int(ny), int(nx)), np.float32)
            optr = vals.ctypes.data_as(ctypes.c_void_p)
        else:
            import torch
            if isinstance(out, str) and out == "device":
                vals = torch.empty((int(nz), int(ny), int(nx)), dtype=torch.float32,
                                   device=torch.device("cuda", self.device))
            else:
                vals = out
            _check_tensor(vals, "out", n, self.device)
            optr = ctypes.c_void_p(vals.data_ptr())
        if out is not None or not isinstance(vkeep, np.ndarray) or not isinstance(fkeep, np.ndarray):
            self._use_torch_stream()  # device buffers: ordered after the caller's work
        st = _lib.DcHipM2sStats()
        self._check(self._lib.dc_hip_mesh_to_sdf(
            self._ctx, vp, n_v, fp, n_f, int(nx), int(ny), int(nz), float(ox), float(oy),
            float(oz), float(vx), float(vy), float(vz), int(bool(voxelize_first)),
            int(voxel_res), optr, ctypes.byref(st)))
        return vals, {f: getattr(st, f) for f, _ in _lib.DcHipM2sStats._fields_}

    def backward(self, grad_vertices, grad_phi):
        """dL/dphi (written into grad_phi: a contiguous float32 CUDA tensor with one element per
        sample of the last forward, all grids' samples concatenated after a batch) for dL/dV
        (3 per vertex of the last forward). The adjoint reads the last forward's samples again:
        the backend keeps the device tensors it was given alive (a contiguous copy, for
        extract_tensor), and they must be unchanged; grad_phi must be a different tensor."""
        import torch
        nv, _ = self.counts()
        if not isinstance(grad_vertices, torch.Tensor):
            raise ValueError("grad_vertices must be a torch tensor")
        gv = grad_vertices.contiguous().to(torch.float32)
        _check_tensor(gv, "grad_vertices", 3 * nv, self.device)
        _check_tensor(grad_phi, "grad_phi", self._samples, self.device)
        self._use_torch_stream()
        st = _lib.DcHipStats()
        self._check(self._lib.dc_hip_backward(self._ctx, ctypes.c_void_p(gv.data_ptr()),
                                              ctypes.c_void_p(grad_phi.data_ptr()),
                                              ctypes.byref(st)))
        return DualContouringStats.from_c(st)


_post_backend = None


def postprocess_mesh(mesh, close_holes=True, remove_floaters=True):
    """The reference's module function postprocess_mesh(mesh, close_holes=True,
    remove_floaters=True) (src/cpp/bindings.cpp:145-146), run on the GPU (device 0): returns a new
    DualContouringMesh with NumPy vertices and faces."""
    global _post_backend
    if _post_backend is None:
        _post_backend = HipDualContouringBackend()
    out, _ = _post_backend.postprocess(mesh, close_holes, remove_floaters)
    return out


_m2s_backend = None


def _m2s(vertices, faces, nx, ny, nz, ox, oy, oz, vx, vy, vz, voxelize_first, voxel_res, out):
    global _m2s_backend
    if _m2s_backend is None:
        _m2s_backend = HipDualContouringBackend()
    vals, st = _m2s_backend.mesh_to_sdf(vertices, faces, nx, ny, nz, ox, oy, oz, vx, vy, vz,
                                        voxelize_first, voxel_res, out)
    return vals, st["total_ms"]


def mesh_to_sdf_hip(vertices, faces, nx, ny, nz, ox, oy, oz, vx, vy, vz, voxelize_first=False,
                    voxel_res=256):
    """The reference module's mesh_to_sdf_cuda (src/cpp/bindings.cpp:270-276) on the GPU (device
    0): (DenseSdfGrid with NumPy values, ms)."""
    vals, ms = _m2s(vertices, faces, nx, ny, nz, ox, oy, oz, vx, vy, vz, voxelize_first,
                    voxel_res, None)
    return DenseSdfGrid(nx, ny, nz, vx, vy, vz, ox, oy, oz, vals.reshape(-1)), ms


def mesh_to_sdf_device_hip(vertices, faces, nx, ny, nz, ox, oy, oz, vx, vy, vz,
                           voxelize_first=False, voxel_res=256):
    """The reference module's mesh_to_sdf_device_cuda (src/cpp/bindings.cpp:278-284): the grid
    stays on the GPU as a DenseSdfGridDevice whose d_values is a CUDA tensor, ready for
    extract_device; (grid, ms)."""
    vals, ms = _m2s(vertices, faces, nx, ny, nz, ox, oy, oz, vx, vy, vz, voxelize_first,
                    voxel_res, "device")
    return DenseSdfGridDevice(nx, ny, nz, vx, vy, vz, ox, oy, oz, d_values=vals.reshape(-1)), ms


class NormalComputationMode:
    """include/dc_cuda.h:5-9: the normal of a crossing edge's hermite entry in the sparse
    backends."""
    FiniteDifference = 0
    EdgeGradient = 1
    PrecomputedGradient = 2  # no gradient grid reaches the kernels: as FiniteDifference


class HipSparseDualContouringBackend:
    """MI355X sparse brick dual contouring (replaces CudaSparseDualContouringBackend,
    include/dc_cuda.h:37-56, bound at src/cpp/bindings.cpp:229-243). extract(grid) and
    extract_device(grid) return (mesh, stats) equal to the reference's: cells of active bricks
    (bricks of brick_size cells whose samples hold both signs) with both signs or a corner within
    0.45 vx of zero, one FP32 QEF vertex per cell (two for a split cell when multi_vertex_cells),
    quads on strict sign changes and both windings on thin edges. chunk_size is kept as an
    attribute and has no effect, as in the reference."""

    _name = "hip-sparse"

    def __init__(self, brick_size=8, normal_mode=NormalComputationMode.FiniteDifference,
                 chunk_size=0, multi_vertex_cells=False, device=0):
        self.brick_size = brick_size
        self.normal_mode = normal_mode
        self.chunk_size = chunk_size
        self.multi_vertex_cells = multi_vertex_cells
        self.device = device
        self._be = HipDualContouringBackend(device=device)  # owns the context

    def _stats(self, st, download_ms):
        out = DualContouringStats()
        out.backend = "hip-sparse-mvdc" if self.multi_vertex_cells else "hip-sparse"
        for f in ("nx", "ny", "nz") + _lib._SPARSE_INT + _lib._SPARSE_MS:
            if hasattr(out, f) or f in ("sample_pass_ms",):
                setattr(out, f, getattr(st, f))
        out.download_ms = download_ms
        out.total_ms = st.total_ms + download_ms
        return out

    def result_counts(self):
        nv, nf = ctypes.c_int64(), ctypes.c_int64()
        self._be._check(self._be._lib.dc_hip_sparse_result_device(
            self._be._ctx, ctypes.byref(nv), ctypes.byref(nf), None, None))
        return nv.value, nf.value

    def _download(self, st):
        import time
        t0 = time.perf_counter()
        nv, nf = self.result_counts()
        V = np.empty(3 * nv, np.float32)
        F = np.empty(4 * nf, np.int32)
        self._be._check(self._be._lib.dc_hip_sparse_copy_result(
            self._be._ctx, V.ctypes.data_as(ctypes.c_void_p), F.ctypes.data_as(ctypes.c_void_p)))
        return DualContouringMesh(V, F), self._stats(st, (time.perf_counter() - t0) * 1e3)

    def extract(self, grid):
        """Host grid -> (DualContouringMesh, DualContouringStats) (dc_cuda_sparse.cu:1134-1168)."""
        vals = np.ascontiguousarray(np.asarray(grid.values, dtype=np.float32))
        if vals.size != int(grid.nx) * int(grid.ny) * int(grid.nz):
            raise ValueError("values size does not match nx*ny*nz")
        cg = self._be._cgrid(grid, vals.ctypes.data)
        st = _lib.DcHipSparseStats()
        self._be.set_stream(None)
        self._be._check(self._be._lib.dc_hip_extract_sparse(
            self._be._ctx, ctypes.byref(cg), int(self.brick_size), int(self.normal_mode),
            int(bool(self.multi_vertex_cells)), ctypes.byref(st)))
        return self._download(st)

    def extract_device(self, grid, download=True):
        """Device grid -> (mesh, stats) (dc_cuda_sparse.cu:1170-1485). grid: DenseSdfGridDevice
        whose d_values is a float32 CUDA tensor (checked) or an int device pointer. With
        download=False the mesh stays on the device: (V [n,3] float32, F [m,4] int32) torch
        tensors copied from the context's result, and the stats."""
        n = int(grid.nx) * int(grid.ny) * int(grid.nz)
        if not isinstance(grid.d_values, int):
            _check_tensor(grid.d_values, "d_values", n, self.device)
            self._be._use_torch_stream()
        cg = self._be._cgrid(grid, _ptr(grid.d_values))
        st = _lib.DcHipSparseStats()
        self._be._check(self._be._lib.dc_hip_extract_sparse_device(
            self._be._ctx, ctypes.byref(cg), int(self.brick_size), int(self.normal_mode),
            int(bool(self.multi_vertex_cells)), ctypes.byref(st)))
        if download:
            return self._download(st)
        import torch
        nv, nf = self.result_counts()
        dev = torch.device("cuda", self.device)
        V = torch.empty((nv, 3), dtype=torch.float32, device=dev)
        F = torch.empty((nf, 4), dtype=torch.int32, device=dev)
        self._be._check(self._be._lib.dc_hip_sparse_copy_result(
            self._be._ctx, ctypes.c_void_p(V.data_ptr()), ctypes.c_void_p(F.data_ptr())))
        return (V, F), self._stats(st, 0.0)


class HipSparseMvdcDualContouringBackend(HipSparseDualContouringBackend):
    """The multi-vertex form (replaces CudaSparseMvdcDualContouringBackend,
    include/dc_cuda.h:58-62, bound at src/cpp/bindings.cpp:245-262)."""

    def __init__(self, brick_size=8, normal_mode=NormalComputationMode.FiniteDifference,
                 chunk_size=0, device=0):
        super().__init__(brick_size, normal_mode, chunk_size, True, device)
