"""ctypes binding of libdc_hip.so (include/dc_hip.h) and libdc_synth.so.

The product path has no fallback: if the HIP library is missing, loading raises.
"""
import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
BUILD_DIR = os.path.join(os.path.dirname(_PKG), "build")

DC_HIP_OK = 0
DC_HIP_ERR_NO_DEVICE = 1
DC_HIP_ERR_INVALID_ARG = 2
DC_HIP_ERR_INVALID_GRID = 3
DC_HIP_ERR_NONFINITE = 4
DC_HIP_ERR_OUT_OF_MEMORY = 5
DC_HIP_ERR_HIP = 6
DC_HIP_ERR_NO_FORWARD = 7
DC_HIP_ERR_INTERNAL = 8

DC_HIP_PLAIN_FP64 = 0
DC_HIP_FP32_FALLBACK = 1

# Every symbol include/dc_hip.h declares; the CPU tests check the library exports all of them.
EXPORTS = (
    "dc_hip_create", "dc_hip_destroy", "dc_hip_set_stream", "dc_hip_set_option",
    "dc_hip_extract", "dc_hip_extract_device", "dc_hip_extract_device_noalloc",
    "dc_hip_result_counts", "dc_hip_copy_result",
    "dc_hip_result_device", "dc_hip_extract_batch", "dc_hip_backward", "dc_hip_autotune",
    "dc_hip_triangulate", "dc_hip_postprocess", "dc_hip_postprocess_copy",
    "dc_hip_postprocess_device", "dc_hip_mesh_to_sdf", "dc_hip_malloc", "dc_hip_free",
    "dc_hip_extract_sparse", "dc_hip_extract_sparse_device", "dc_hip_sparse_result_device",
    "dc_hip_sparse_copy_result", "dc_hip_last_error", "dc_hip_status_string",
)


class DcHipGrid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("vx", ctypes.c_float), ("vy", ctypes.c_float), ("vz", ctypes.c_float),
                ("ox", ctypes.c_float), ("oy", ctypes.c_float), ("oz", ctypes.c_float),
                ("values", ctypes.c_void_p)]


class DcHipStats(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("total_cells", ctypes.c_int64), ("active_cells", ctypes.c_int64),
                ("upload_ms", ctypes.c_float), ("marking_ms", ctypes.c_float),
                ("compaction_ms", ctypes.c_float), ("face_prefix_sum_ms", ctypes.c_float),
                ("qef_ms", ctypes.c_float), ("face_fill_ms", ctypes.c_float),
                ("face_emission_ms", ctypes.c_float), ("download_ms", ctypes.c_float),
                ("total_ms", ctypes.c_float), ("backward_ms", ctypes.c_float),
                ("vertex_count", ctypes.c_int64), ("face_count", ctypes.c_int64),
                ("qef_fallback_count", ctypes.c_int64), ("invalid_count", ctypes.c_int64),
                ("normals_ms", ctypes.c_float), ("mark_kernel_ms", ctypes.c_float),
                ("graph_mode", ctypes.c_int32)]


class DcHipPostStats(ctypes.Structure):
    _fields_ = [("welded_vertices", ctypes.c_int64), ("welded_faces", ctypes.c_int64),
                ("components", ctypes.c_int64), ("kept_faces", ctypes.c_int64),
                ("boundary_edges", ctypes.c_int64), ("loops_filled", ctypes.c_int64),
                ("added_vertices", ctypes.c_int64), ("added_faces", ctypes.c_int64),
                ("vertex_count", ctypes.c_int64), ("face_count", ctypes.c_int64),
                ("walk_ms", ctypes.c_float), ("total_ms", ctypes.c_float)]


class DcHipM2sStats(ctypes.Structure):
    _fields_ = [("triangles", ctypes.c_int64), ("bin_entries", ctypes.c_int64),
                ("bins", ctypes.c_int64), ("active_bins", ctypes.c_int64),
                ("flood_passes", ctypes.c_int32), ("voxel_res", ctypes.c_int32),
                ("eval_ms", ctypes.c_float), ("total_ms", ctypes.c_float)]


_SPARSE_INT = ("total_cells", "active_cells", "total_bricks", "active_bricks", "vertex_count",
               "face_count", "qef_fallback_count", "clamp_count", "invalid_count",
               "ambiguous_cells", "multi_vertex_cells", "one_vertex_cells", "two_vertex_cells",
               "split_rejection_count", "bad_qef_count",
               "faces_skipped_due_to_missing_cluster")
_SPARSE_MS = ("upload_ms", "active_brick_marking_ms", "active_brick_compaction_ms",
              "active_cell_marking_ms", "active_cell_compaction_ms", "marking_ms",
              "compaction_ms", "qef_ms", "face_count_ms", "face_prefix_sum_ms", "face_fill_ms",
              "face_emission_ms", "download_ms", "total_ms", "sample_pass_ms")


class DcHipSparseStats(ctypes.Structure):
    _fields_ = ([("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                 ("brick_size", ctypes.c_int32)]
                + [(f, ctypes.c_int64) for f in _SPARSE_INT]
                + [(f, ctypes.c_float) for f in _SPARSE_MS])


_hip = None
_synth = None


def lib_path():
    return os.path.join(BUILD_DIR, "libdc_hip.so")


def load_hip():
    """Load libdc_hip.so; raises RuntimeError when the HIP extension has not been built."""
    global _hip
    if _hip is not None:
        return _hip
    path = lib_path()
    if not os.path.exists(path):
        raise RuntimeError(f"HIP extension not built: {path} is missing (run __graft_entry__.build())")
    # One HIP runtime per process: torch ships its own libamdhip64.so.7 and loads it by path. Once
    # it is loaded, the library's dependency on that soname resolves to it; loaded first, the
    # library would bring in the system runtime beside it, and whichever of the two initialized
    # the device second would find no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    P, I, I64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
    sig = {
        "dc_hip_create": (I, [I, ctypes.POINTER(P)]),
        "dc_hip_destroy": (None, [P]),
        "dc_hip_set_stream": (I, [P, P]),
        "dc_hip_set_option": (I, [P, ctypes.c_char_p, I64]),
        "dc_hip_extract": (I, [P, ctypes.POINTER(DcHipGrid), I, ctypes.POINTER(DcHipStats)]),
        "dc_hip_extract_device": (I, [P, ctypes.POINTER(DcHipGrid), I,
                                      ctypes.POINTER(DcHipStats)]),
        "dc_hip_extract_device_noalloc": (I, [P, ctypes.POINTER(DcHipGrid), I, P, P, I64, P, I64,
                                              ctypes.POINTER(I64), ctypes.POINTER(I64)]),
        "dc_hip_result_counts": (I, [P, ctypes.POINTER(I64), ctypes.POINTER(I64)]),
        "dc_hip_copy_result": (I, [P, P, P, P]),
        "dc_hip_result_device": (I, [P, ctypes.POINTER(P), ctypes.POINTER(P), ctypes.POINTER(P)]),
        "dc_hip_extract_batch": (I, [P, ctypes.POINTER(DcHipGrid), ctypes.c_int32, I, P, P,
                                     ctypes.POINTER(DcHipStats)]),
        "dc_hip_backward": (I, [P, P, P, ctypes.POINTER(DcHipStats)]),
        "dc_hip_triangulate": (I, [P, P]),
        "dc_hip_autotune": (I, [P, ctypes.POINTER(DcHipGrid), ctypes.c_int32, I, ctypes.c_int32,
                                ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                ctypes.POINTER(ctypes.c_float)]),
        "dc_hip_postprocess": (I, [P, P, I64, P, I64, I, I, ctypes.POINTER(I64),
                                   ctypes.POINTER(I64), ctypes.POINTER(DcHipPostStats)]),
        "dc_hip_postprocess_copy": (I, [P, P, P]),
        "dc_hip_postprocess_device": (I, [P, ctypes.POINTER(P), ctypes.POINTER(P)]),
        "dc_hip_mesh_to_sdf": (I, [P, P, I64, P, I64, ctypes.c_int32, ctypes.c_int32,
                                   ctypes.c_int32] + [ctypes.c_float] * 6
                               + [I, ctypes.c_int32, P, ctypes.POINTER(DcHipM2sStats)]),
        "dc_hip_malloc": (I, [P, I64, ctypes.POINTER(P)]),
        "dc_hip_extract_sparse": (I, [P, ctypes.POINTER(DcHipGrid), ctypes.c_int32, I, I,
                                      ctypes.POINTER(DcHipSparseStats)]),
        "dc_hip_extract_sparse_device": (I, [P, ctypes.POINTER(DcHipGrid), ctypes.c_int32, I, I,
                                             ctypes.POINTER(DcHipSparseStats)]),
        "dc_hip_sparse_result_device": (I, [P, ctypes.POINTER(I64), ctypes.POINTER(I64),
                                            ctypes.POINTER(P), ctypes.POINTER(P)]),
        "dc_hip_sparse_copy_result": (I, [P, P, P]),
        "dc_hip_free": (I, [P]),
        "dc_hip_last_error": (ctypes.c_char_p, [P]),
        "dc_hip_status_string": (ctypes.c_char_p, [I]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _hip = lib
    return lib


def load_synth():
    global _synth
    if _synth is not None:
        return _synth
    path = os.path.join(BUILD_DIR, "libdc_synth.so")
    if not os.path.exists(path):
        raise RuntimeError(f"synth library not built: {path} is missing")
    lib = ctypes.CDLL(path)
    lib.dc_synth_csg.restype = ctypes.c_int
    lib.dc_synth_csg.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
    lib.dc_synth_cli.restype = ctypes.c_int
    lib.dc_synth_cli.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    _synth = lib
    return lib


class DcHipError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"dc_hip error {status}: {message}")
        self.status = status
