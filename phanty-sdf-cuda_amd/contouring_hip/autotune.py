"""Autotuner for the marking sign pass's tile shape (sample rows per wave tile, waves per block).

The reference CUDA backend launches its marking kernel with one fixed block shape
(src/cuda/dc_cuda.cu:687-692). Here the shape is picked per grid size by the library's own sweep
(dc_hip_autotune: every candidate timed on the caller's grids with the pipeline's hipEvent pass
timers), and the winner is cached on disk keyed by device name and grid shape, so later processes
skip the sweep.
"""
import ctypes
import itertools
import json
import os

from . import _lib

ROWS = (4, 8, 16, 32)
WAVES = (1, 2, 4)


def _default_cache():
    base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return os.path.join(base, "phanty_dc_hip", "autotune.json")


def _device_name(device):
    try:
        import torch
        return torch.cuda.get_device_name(device)
    except Exception:
        return f"device{device}"


def _key(device, grids):
    shapes = ",".join(f"{int(g.nx)}x{int(g.ny)}x{int(g.nz)}" for g in grids[:4])
    total = sum(int(g.nx) * int(g.ny) * int(g.nz) for g in grids)
    return f"rows|{_device_name(device)}|{len(grids)}|{shapes}|{total}"


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path, table):
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + f".{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # the cache is an optimization only


def _apply(backend, cfg):
    backend.set_option("mark_rows", cfg["rows"])
    backend.set_option("mark_waves", cfg["waves"])


def autotune_mark(backend, grids, repeats=3, cache_path=None, use_cache=True):
    """Pick the fastest mark tile shape for `grids` (DenseSdfGridDevice or a list of them, values
    resident on the backend's device) and set it on `backend`.

    Returns (best, table): best = {"rows", "waves", "mark_ms"}; table = every candidate
    timed (empty when the cached choice was used).
    """
    if not isinstance(grids, (list, tuple)):
        grids = [grids]
    cache_path = cache_path or _default_cache()
    key = _key(backend.device, grids)
    if use_cache:
        hit = _load(cache_path).get(key)
        if hit:
            _apply(backend, hit)
            return hit, []
    arr = (_lib.DcHipGrid * len(grids))(
        *[backend._cgrid(g, g.d_values if isinstance(g.d_values, int) else g.d_values.data_ptr())
          for g in grids])
    from .backend import PRECISIONS
    prec = PRECISIONS[backend.precision]
    if not all(isinstance(g.d_values, int) for g in grids):
        backend._use_torch_stream()
    ms = (ctypes.c_float * (len(ROWS) * len(WAVES)))()
    bl, bw = ctypes.c_int32(), ctypes.c_int32()
    backend._check(backend._lib.dc_hip_autotune(backend._ctx, arr, len(grids), prec, repeats,
                                                ctypes.byref(bl), ctypes.byref(bw), ms))
    table = [{"rows": rows, "waves": waves, "mark_ms": float(ms[q])}
             for q, (rows, waves) in enumerate(itertools.product(ROWS, WAVES))]
    table.sort(key=lambda c: c["mark_ms"])
    winner = {"rows": int(bl.value), "waves": int(bw.value),
              "mark_ms": next(c["mark_ms"] for c in table
                              if (c["rows"], c["waves"]) == (bl.value, bw.value))}
    if use_cache:
        t = _load(cache_path)
        t[key] = winner
        _store(cache_path, t)
    return winner, table
