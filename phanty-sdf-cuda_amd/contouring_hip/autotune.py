"""Autotuners: the marking sign pass's tile shape (sample rows per wave tile, waves per block), and
the per-cell kernels' launch shapes (solve blocks per CU, gather rows per wave).

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


# Launch shapes of the per-cell kernels, chosen by whole timed steps (after the mark tile): the
# solve kernels' blocks per CU (option qef_blocks: the FP64 solve's grid-stride loop, the FP32
# solve's contiguous shares and twice as many for the FP64 fallback; 0: the library's default, 8
# per CU), and, for the adjoint, sample rows per wave of the gradient gather. Neither changes the mesh; the gather's
# row count may change the order of each sample's (at most 8) corner terms, within 1e-6 relative.
QEF_BLOCKS_PER_CU = (0, 4, 16, 32)
GATHER_ROWS = (1, 2, 4)


def _cu_count(device):
    try:
        import torch
        return int(torch.cuda.get_device_properties(device).multi_processor_count)
    except Exception:
        return 256


def _median_step_ms(step, repeats, warmup=2, per=5):
    import time
    import torch
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    samples = []
    for _ in range(max(repeats, 1)):
        t0 = time.perf_counter()
        for _ in range(per):
            step()
        torch.cuda.synchronize()
        samples.append((time.perf_counter() - t0) * 1e3 / per)
    samples.sort()
    return samples[len(samples) // 2]


def autotune_launch(backend, grid, repeats=5, adjoint=False, cache_path=None, use_cache=True):
    """Pick the per-cell kernels' launch shapes for `grid` (a DenseSdfGridDevice: values as a
    float32 CUDA tensor on the backend's device, or a device pointer) by timing whole extractions (and, with adjoint=True,
    extraction + adjoint steps) with the launch sequences replayed, and set them on `backend`.

    Returns (best, table): best = {"qef_blocks", "gather_rows", "step_ms"} (qef_blocks in blocks of
    256 threads, 0 = the library's default); table = every candidate timed (empty when the cached
    choice was used). Per-pass timing is switched off while timing and restored afterwards.
    """
    import torch
    from .backend import PRECISIONS, _check_tensor, _ptr
    cache_path = cache_path or _default_cache()
    key = "launch|" + _key(backend.device, [grid]) + f"|{backend.precision}|{int(bool(adjoint))}"
    if use_cache:
        hit = _load(cache_path).get(key)
        if hit:
            backend.set_option("qef_blocks", hit["qef_blocks"])
            backend.set_option("gather_rows", hit["gather_rows"])
            return hit, []
    n_samples = int(grid.nx) * int(grid.ny) * int(grid.nz)
    if not isinstance(grid.d_values, int):
        _check_tensor(grid.d_values, "d_values", n_samples, backend.device)
    backend._use_torch_stream()
    dev = torch.device("cuda", backend.device)
    cg = backend._cgrid(grid, _ptr(grid.d_values))
    prec = PRECISIONS[backend.precision]
    lib = backend._lib
    st = _lib.DcHipStats()
    bst = _lib.DcHipStats()
    timing = backend.option("timing", 1)
    backend.set_option("timing", 0)
    state = {}

    def fwd():
        backend._check(lib.dc_hip_extract_device(backend._ctx, ctypes.byref(cg), prec,
                                                 ctypes.byref(st)))

    def step():
        fwd()
        if adjoint:
            if "gv" not in state or state["gv"].shape[0] != int(st.vertex_count):
                gen = torch.Generator(device=dev)
                gen.manual_seed(0)
                state["gv"] = torch.randn((max(int(st.vertex_count), 1), 3), device=dev,
                                          generator=gen)
                state["g"] = torch.empty(n_samples, dtype=torch.float32, device=dev)
            backend._check(lib.dc_hip_backward(
                backend._ctx, ctypes.c_void_p(state["gv"].data_ptr()),
                ctypes.c_void_p(state["g"].data_ptr()), ctypes.byref(bst)))

    cu = _cu_count(backend.device)
    table = []
    try:
        for per_cu in QEF_BLOCKS_PER_CU:
            for rows in (GATHER_ROWS if adjoint else (1,)):
                blocks = per_cu * cu
                backend.set_option("qef_blocks", blocks)
                backend.set_option("gather_rows", rows)
                fwd()  # lays out the gather's row groups for this row count
                table.append({"qef_blocks": blocks, "gather_rows": rows,
                              "step_ms": _median_step_ms(step, repeats)})
    finally:
        backend.set_option("timing", timing)
    table.sort(key=lambda c: c["step_ms"])
    best = dict(table[0])
    backend.set_option("qef_blocks", best["qef_blocks"])
    backend.set_option("gather_rows", best["gather_rows"])
    if use_cache:
        t = _load(cache_path)
        t[key] = best
        _store(cache_path, t)
    return best, table
