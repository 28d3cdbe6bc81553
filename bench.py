"""Benchmark: dense dual contouring of the 512^3 sharp-feature CSG grid on one MI355X.

Workload (one step = one full extraction, grid resident in HBM when timing starts): union of 8
Haar-rotated boxes minus a sphere, sampled as FP32 on a 512^3 unit-cube grid (seed 1234), plain
FP64 mode through the C ABI (dc_hip_extract_device). The mesh stays in device memory; each step
ends with the pipeline's one host synchronization (counts readback).

Prints one JSON line. `value` is ms per extraction (lower is better). `roofline` is the marking
sign pass k_mark (the pass that reads every sample): algorithmic bytes per launch = every sample
read once (4*N_s; its sign-mask output is an intermediate of this design and not counted), over
its hipEvent-measured duration; `marking_stage` prices the whole marking stage (sign and
cell passes) at 4*N_s + N_c/8; `traffic` is its measured HBM bytes per launch from the committed
rocprofv3 counter summary (profiles/k_mark_csg8_512_traffic.json). `cpu_baseline` is the oracle
(the C++ restatement of the reference CPU backend) on the same grid, single thread, on this host.
The mark tile shape is the library default; --tune sweeps the shapes on this grid first.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "phanty-sdf-cuda_amd")]

HBM_PEAK_GBS = 8000.0
PROFILE_DIR = os.path.join(ROOT, "profiles")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse_traffic(n):
    """HBM bytes per mark launch from the committed rocprofv3 PMC summary, if present."""
    path = os.path.join(PROFILE_DIR, f"k_mark_csg8_{n}_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return float(json.load(f)["mark_hbm_bytes_per_launch"])
    except Exception:
        return None


def run_steps(be, lib, cg, prec, steps):
    import ctypes
    from contouring_hip import _lib
    st = _lib.DcHipStats()
    acc = {"marking_ms": 0.0, "mark_kernel_ms": 0.0, "compaction_ms": 0.0,
           "face_prefix_sum_ms": 0.0, "normals_ms": 0.0, "qef_ms": 0.0, "face_fill_ms": 0.0}
    for _ in range(steps):
        rc = lib.dc_hip_extract_device(be._ctx, ctypes.byref(cg), prec, ctypes.byref(st))
        be._check(rc)
        for k in acc:
            acc[k] += getattr(st, k)
    return {k: v / max(steps, 1) for k, v in acc.items()}, st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-runs", type=int, default=5,
                    help="timed single-thread oracle runs (after one warm-up run)")
    ap.add_argument("--no-host-input", action="store_true",
                    help="skip the host-input (upload + download) timing")
    ap.add_argument("--host-steps", type=int, default=5)
    ap.add_argument("--post-runs", type=int, default=10,
                    help="timed post-processing calls (median reported)")
    ap.add_argument("--tune", action="store_true",
                    help="sweep the mark tile shapes on this grid first and use the fastest")
    ap.add_argument("--config", default="c2",
                    choices=("c1", "c2", "c3", "c4", "c5", "m2s", "sparse"),
                    help="c1: 64^3 spheres with oracle parity; "
                         "c2: the 512^3 CSG headline (default); c3: 1024^3 narrow band, plain vs "
                         "FP32+fallback; c4: 256 uneven grids, one batch vs sequential; c5: "
                         "forward+adjoint at 128^3, 256^3 and 512^3; m2s: mesh to SDF of the csg8 DC mesh; "
                         "sparse: the sparse brick backends on csg8")
    ap.add_argument("--calibrate", action="store_true",
                    help="run one elementwise kernel over the grid (known bytes: reads and writes "
                         "4*N_s) for scaling the HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; the dual-contouring "
                                   "path is measured on one MI355X (multi-GPU sharding is a "
                                   "design note, not a run)", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    from contouring_hip import DenseSdfGridDevice, HipDualContouringBackend, _lib, synth

    if args.config != "c2":
        {"c1": bench_c1, "c3": bench_c3, "c4": bench_c4, "c5": bench_c5,
         "m2s": bench_m2s_config, "sparse": bench_sparse_config}[args.config](args)
        return

    n = args.size
    t = time.time()
    g = synth.csg8(n, seed=args.seed)
    log(f"[bench] generated csg8 {n}^3 in {time.time() - t:.1f}s")
    phi = torch.from_numpy(g.values).to("cuda:0")
    be = HipDualContouringBackend("plain-fp64", device=0)
    be.set_stream(torch.cuda.current_stream(0).cuda_stream)
    lib = be._lib
    cg = be._cgrid(g, phi.data_ptr())
    prec = _lib.DC_HIP_PLAIN_FP64

    # mark tile shape: timed on this grid (the sweep is not part of the timed region)
    gdev = DenseSdfGridDevice(g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz, d_values=phi)
    if args.tune:
        tile, table = be.autotune(gdev, repeats=3, use_cache=False)
    else:
        tile, table = {"rows": 16, "waves": 4}, []  # the library's defaults
    if args.calibrate:
        # a kernel of known traffic: its FETCH_SIZE / WRITE_SIZE scale the marking counters
        dst = torch.empty_like(phi)
        torch.mul(phi, 1.0, out=dst)
        torch.cuda.synchronize()
        del dst
    for row in table[:6] + table[-2:]:
        log(f"[tune] rows {row['rows']} waves {row['waves']}: mark {row['mark_ms'] * 1e3:.1f} us")

    # per-pass device times (events between the kernels: eager launches), after a warm-up that
    # sizes the buffers
    run_steps(be, lib, cg, prec, max(args.warmup, 1))
    passes, st = run_steps(be, lib, cg, prec, max(3, args.steps // 2))
    # the timed region: no per-pass events, so the launch sequence is captured once and replayed
    be.set_option("timing", 0)
    run_steps(be, lib, cg, prec, max(args.warmup, 2))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, st = run_steps(be, lib, cg, prec, args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    graph_mode = int(st.graph_mode)
    log(f"[bench] plain fp64: {ms:.3f} ms/extraction (launch sequence "
        f"{'replayed' if graph_mode == 2 else 'eager'}), passes {passes}")

    # plain-mode FP64 vertices (bit-identical to the CPU backend's) for the FP32 error below
    nv = int(st.vertex_count)
    nf = int(st.face_count)
    v64 = torch.empty((nv, 3), dtype=torch.float64, device="cuda:0")
    f_plain = torch.empty((nf, 4), dtype=torch.int32, device="cuda:0")
    be._check(lib.dc_hip_copy_result(be._ctx, None, ctypes.c_void_p(v64.data_ptr()),
                                     ctypes.c_void_p(f_plain.data_ptr())))

    # FP32 + FP64-fallback forward followed by the adjoint (dL/dphi for a fixed dL/dV)
    prec32 = _lib.DC_HIP_FP32_FALLBACK
    gen = torch.Generator(device="cuda:0")
    gen.manual_seed(0)
    gv = torch.randn((nv, 3), dtype=torch.float32, device="cuda:0", generator=gen)
    gphi = torch.empty_like(phi)
    bst = _lib.DcHipStats()

    def fwd_bwd(k):
        acc = {"fwd_ms": 0.0, "bwd_ms": 0.0}
        for _ in range(k):
            rc = lib.dc_hip_extract_device(be._ctx, ctypes.byref(cg), prec32, ctypes.byref(st))
            be._check(rc)
            rc = lib.dc_hip_backward(be._ctx, ctypes.c_void_p(gv.data_ptr()),
                                     ctypes.c_void_p(gphi.data_ptr()), ctypes.byref(bst))
            be._check(rc)
        return acc

    fwd_bwd(max(args.warmup, 2))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fwd_bwd(args.steps)
    torch.cuda.synchronize()
    fb_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    be.set_option("timing", 1)  # one more step with events for the adjoint's own time
    fwd_bwd(1)
    fallback = int(st.qef_fallback_count)
    v32 = torch.empty((nv, 3), dtype=torch.float32, device="cuda:0")
    f32 = torch.empty((int(st.face_count), 4), dtype=torch.int32, device="cuda:0")
    be._check(lib.dc_hip_copy_result(be._ctx, ctypes.c_void_p(v32.data_ptr()), None,
                                     ctypes.c_void_p(f32.data_ptr())))
    err = (v32.double() - v64).abs() / g.h
    err_max, err_rms = float(err.max()), float(err.pow(2).mean().sqrt())
    # quads whose orientation differs from the plain (CPU-exact) mesh
    flips = int((f32 != f_plain).any(dim=1).sum()) if f32.shape == f_plain.shape else -1
    log(f"[bench] fp32+fallback fwd+bwd: {fb_ms:.3f} ms ({fb_ms / ms:.2f}x plain), fallback "
        f"{fallback}/{nv}, vertex err max {err_max:.2e} h rms {err_rms:.2e} h, "
        f"bwd {bst.backward_ms:.3f} ms, quads oriented unlike plain: {flips}")

    post = bench_postprocess(be, lib, cg, prec, st, args)
    m2s = bench_mesh_to_sdf(be, lib, cg, prec, st, args)
    sparse = bench_sparse(g, phi, args)

    Ns = n ** 3
    Nc = (n - 1) ** 3
    # sign pass: the algorithmic bytes are the samples it reads (4 N_s); its 16-byte sign entries
    # per pencil and 64 layers are this design's intermediate, not counted
    mark_bytes = 4 * Ns
    mark_s = passes["mark_kernel_ms"] * 1e-3
    achieved = mark_bytes / mark_s / 1e9 if mark_s > 0 else None
    stage_bytes = 4 * Ns + Nc / 8
    stage_s = passes["marking_ms"] * 1e-3
    stage_gbs = stage_bytes / stage_s / 1e9 if stage_s > 0 else None
    traffic = parse_traffic(n)

    cpu = None
    if not args.no_cpu:
        import oracle
        host = host_cpu()
        # the reference CPU backend is single-threaded: one thread, median of cpu_runs after a
        # warm-up run
        oracle.set_threads(1)
        times = []
        for q in range(args.cpu_runs + 1):
            t = time.perf_counter()
            r = oracle.run_grid(g)
            if q > 0:
                times.append((time.perf_counter() - t) * 1e3)
            del r
            log(f"[bench] oracle run {q} (1 thread){' warm-up' if q == 0 else ''}")
        cpu_ms = float(np.median(times))
        # the same restatement over all host cores (its loops are OpenMP), reported beside it
        oracle.set_threads(host["nproc"])
        t = time.perf_counter()
        r = oracle.run_grid(g)
        all_ms = (time.perf_counter() - t) * 1e3
        ok = (r.n_vertices == st.vertex_count and r.n_faces == st.face_count)
        del r
        log(f"[bench] oracle: {cpu_ms:.0f} ms (1 thread, median of {len(times)}), "
            f"{all_ms:.0f} ms ({host['nproc']} threads), counts match GPU: {ok}")
        cpu = {"value": cpu_ms, "unit": "ms per extraction", "cores": 1, "kind": "port",
               "runs_ms": times, "host_nproc": host["nproc"], "host_cpu": host["model"],
               "all_cores_ms": all_ms,
               "sample": f"full csg8 {n}^3 extractions (seed {args.seed}) by the C++ "
                         f"restatement of CpuDualContouringBackend: median of {len(times)} "
                         "single-thread runs after one warm-up run"}

    host_input = None
    if not args.no_host_input:
        host_input = time_host_input(be, g, args)

    out = {
        "metric": "512^3 DC ms (plain vs FP32+bwd, <=1e-4*h vertex err); %HBM peak; CPU ms",
        "value": ms,
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"csg8 sharp-feature CSG {n}^3 FP32 grid, plain FP64 solve",
                   "grid": n, "seed": args.seed, "precision": "plain-fp64",
                   "mark_tile": {k: tile[k] for k in ("rows", "waves")},
                   "parallelism": "single GPU"},
        "active_cells": int(st.vertex_count),
        "faces": int(st.face_count),
        "active_cells_per_s": st.vertex_count / (ms * 1e-3),
        "pass_ms": passes,
        "launch": "hipGraph replay" if graph_mode == 2 else "eager",
        "fp32_bwd_ms": fb_ms,
        "fp32_bwd_over_plain": fb_ms / ms,
        "fp32_vertex_err_h": {"max": err_max, "rms": err_rms},
        "fp32_fallback_cells": fallback,
        "fp32_orientation_flips": flips,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": traffic, "kernel": "k_mark",
                     "algorithmic_bytes": mark_bytes},
        "marking_stage": {"kernels": ["k_mark", "k_cells"], "ms": passes["marking_ms"],
                          "algorithmic_bytes": stage_bytes, "achieved_gbs": stage_gbs,
                          "frac": (stage_gbs / HBM_PEAK_GBS) if stage_gbs else None},
        "cpu_baseline": cpu,
        "host_input": host_input,
        "postprocess": post,
        "mesh_to_sdf": m2s,
        "sparse": sparse,
    }
    print(json.dumps(out), flush=True)


def bench_postprocess(be, lib, cg, prec, st, args):
    """postprocess_mesh (weld, floaters, holes) of the plain extraction's mesh on the device: host
    wall time per call (median; it synchronises for its counts and the boundary walk), beside the
    oracle's single-thread restatement on the same mesh (skipped with --no-cpu)."""
    be._check(lib.dc_hip_extract_device(be._ctx, ctypes.byref(cg), prec, ctypes.byref(st)))
    times = []
    for q in range(args.post_runs + 2):
        t = time.perf_counter()
        out, pst = be.postprocess(None, True, True)
        if q >= 2:
            times.append((time.perf_counter() - t) * 1e3)
    rec = {"gpu_ms": float(np.median(times)), "runs": len(times), "close_holes": True,
           "remove_floaters": True, "faces_in": int(st.face_count),
           "faces_out": int(pst["face_count"]), "walk_ms": float(pst["walk_ms"]),
           "boundary_edges": int(pst["boundary_edges"]), "loops_filled": int(pst["loops_filled"])}
    log(f"[bench] postprocess: {rec['gpu_ms']:.2f} ms (median of {len(times)}), "
        f"{rec['faces_in']} -> {rec['faces_out']} quads")
    if not args.no_cpu:
        import oracle
        mesh = be._download(False)
        t = time.perf_counter()
        V2, F2, _ = oracle.postprocess(mesh.vertices.reshape(-1, 3), mesh.faces.reshape(-1, 4))
        rec["cpu_ms"] = (time.perf_counter() - t) * 1e3
        rec["cpu_kind"] = "port (oracle/post_oracle.cpp, 1 thread)"
        rec["equal_to_oracle"] = bool(np.array_equal(V2.ravel(), out.vertices) and
                                      np.array_equal(F2.ravel(), out.faces))
        log(f"[bench] postprocess oracle: {rec['cpu_ms']:.0f} ms, equal: {rec['equal_to_oracle']}")
    return rec


def bench_mesh_to_sdf(be, lib, cg, prec, st, args, cpu_n=32):
    """The mesh-to-SDF pass (dc_hip_mesh_to_sdf) on the plain extraction's mesh (quads split
    into triangles) sampled back on the same grid: host wall time per call (median of 5 after a
    warm-up; the output stays on the device), beside the oracle's single-thread restatement and
    the GPU on a small csg8 grid (cpu_n^3, its own DC mesh) where the oracle finishes in seconds
    (skipped with --no-cpu)."""
    import torch
    from contouring_hip import DenseSdfGrid, synth
    be._check(lib.dc_hip_extract_device(be._ctx, ctypes.byref(cg), prec, ctypes.byref(st)))
    mesh = be._download(False)
    V = torch.from_numpy(mesh.vertices.reshape(-1, 3)).cuda()
    F = torch.from_numpy(synth.quads_to_triangles(mesh.faces)).cuda()
    n = cg.nx
    out = torch.empty(cg.nx * cg.ny * cg.nz, dtype=torch.float32, device="cuda:0")
    times = []
    for q in range(6):
        t = time.perf_counter()
        _, mst = be.mesh_to_sdf(V, F, cg.nx, cg.ny, cg.nz, cg.ox, cg.oy, cg.oz, cg.vx, cg.vy,
                                cg.vz, out=out)
        if q > 0:
            times.append((time.perf_counter() - t) * 1e3)
    rec = {"workload": f"csg8 {n}^3 DC mesh ({int(F.shape[0])} triangles) sampled on its "
                       f"{n}^3 grid", "gpu_ms": float(np.median(times)), "runs": len(times),
           "eval_ms": float(mst["eval_ms"]), "triangles": int(mst["triangles"]),
           "bin_entries": int(mst["bin_entries"]), "active_bins": int(mst["active_bins"]),
           "flood_passes": int(mst["flood_passes"]),
           "samples_per_s": n ** 3 / (float(np.median(times)) * 1e-3)}
    log(f"[bench] mesh_to_sdf {n}^3: {rec['gpu_ms']:.2f} ms (evaluation {rec['eval_ms']:.2f} ms)")
    if not args.no_cpu:
        import oracle
        gs = synth.csg8(cpu_n, seed=args.seed)
        ms_, _ = be.extract(DenseSdfGrid(gs.nx, gs.ny, gs.nz, gs.vx, gs.vy, gs.vz, gs.ox, gs.oy,
                                         gs.oz, gs.values))
        Vs = ms_.vertices.reshape(-1, 3)
        Fs = synth.quads_to_triangles(ms_.faces)
        dims, origin, spacing = (cpu_n,) * 3, (gs.ox, gs.oy, gs.oz), (gs.vx, gs.vy, gs.vz)
        oracle.set_threads(1)
        t = time.perf_counter()
        want = oracle.mesh_to_sdf(Vs, Fs, dims, origin, spacing)
        cpu_ms = (time.perf_counter() - t) * 1e3
        gt = []
        for q in range(6):
            t = time.perf_counter()
            got, _ = be.mesh_to_sdf(Vs, Fs, *dims, *origin, *spacing)
            if q > 0:
                gt.append((time.perf_counter() - t) * 1e3)
        rec["small"] = {"grid": cpu_n, "triangles": int(Fs.shape[0]), "cpu_ms": cpu_ms,
                        "cpu_kind": "port (oracle/m2s_oracle.cpp, 1 thread)",
                        "gpu_ms": float(np.median(gt)),
                        "equal_to_oracle": bool(np.array_equal(got.view(np.uint32),
                                                               want.view(np.uint32)))}
        log(f"[bench] mesh_to_sdf {cpu_n}^3: oracle {cpu_ms:.0f} ms (1 thread), GPU "
            f"{rec['small']['gpu_ms']:.2f} ms, equal: {rec['small']['equal_to_oracle']}")
    return rec


def bench_sparse(g, phi, args, cpu_n=128):
    """The sparse brick backends (dc_hip_extract_sparse_device, brick 8, finite-difference
    normals) on the same device grid: host wall time per call with the mesh left on the device
    (median of 5 after a warm-up; each call synchronises for its counts), the per-stage device
    times and the sample pass's read rate (4 bytes per sample), for the one-vertex and the
    multi-vertex form; beside the oracle's single-thread restatement and the GPU on a csg8 grid of
    cpu_n^3 (skipped with --no-cpu)."""
    from contouring_hip import (DenseSdfGridDevice, HipSparseDualContouringBackend,
                                HipSparseMvdcDualContouringBackend, _lib, synth)
    n = g.nx
    recs = {}
    for mv in (False, True):
        sp = HipSparseMvdcDualContouringBackend() if mv else HipSparseDualContouringBackend()
        be = sp._be
        be._use_torch_stream()
        cg = be._cgrid(g, phi.data_ptr())
        st = _lib.DcHipSparseStats()
        times = []
        for q in range(6):
            t = time.perf_counter()
            be._check(be._lib.dc_hip_extract_sparse_device(be._ctx, ctypes.byref(cg), 8, 0,
                                                           int(mv), ctypes.byref(st)))
            if q > 0:
                times.append((time.perf_counter() - t) * 1e3)
        ms = float(np.median(times))
        rec = {"workload": f"csg8 {n}^3, brick 8, finite-difference normals", "gpu_ms": ms,
               "runs": len(times), "active_cells_per_s": st.active_cells / (ms * 1e-3),
               "sample_pass_gbs": 4.0 * n ** 3 / (st.sample_pass_ms * 1e-3) / 1e9
               if st.sample_pass_ms > 0 else None}
        for f in _lib._SPARSE_INT + _lib._SPARSE_MS:
            rec[f] = getattr(st, f)
        log(f"[bench] {'mvdc' if mv else 'sparse'} {n}^3: {ms:.3f} ms, {st.active_bricks}/"
            f"{st.total_bricks} bricks, {st.active_cells} cells, {st.vertex_count} vertices, "
            f"{st.face_count} quads; sample pass {st.sample_pass_ms * 1e3:.1f} us "
            f"({rec['sample_pass_gbs']:.0f} GB/s), cells {st.active_cell_marking_ms * 1e3:.1f} us, "
            f"scan {st.active_cell_compaction_ms * 1e3:.1f} us, vertices {st.qef_ms * 1e3:.1f} us, "
            f"faces {st.face_emission_ms * 1e3:.1f} us")
        if not args.no_cpu:
            import oracle
            gs = synth.csg8(cpu_n, seed=args.seed)
            t = time.perf_counter()
            want = oracle.sparse_grid(gs, multi_vertex=mv)
            cpu_ms = (time.perf_counter() - t) * 1e3
            mesh, _ = sp.extract(gs)
            gt = []
            for q in range(6):
                t = time.perf_counter()
                mesh, _ = sp.extract(gs)
                if q > 0:
                    gt.append((time.perf_counter() - t) * 1e3)
            rec["small"] = {
                "grid": cpu_n, "cpu_ms": cpu_ms,
                "cpu_kind": "port (oracle/sparse_oracle.cpp, 1 thread)",
                "gpu_ms_host_grid": float(np.median(gt)),
                "equal_to_oracle": bool(
                    np.array_equal(mesh.vertices.view(np.uint32),
                                   want["vertices"].ravel().view(np.uint32)) and
                    np.array_equal(mesh.faces, want["faces"].ravel()))}
            log(f"[bench] {'mvdc' if mv else 'sparse'} {cpu_n}^3: oracle {cpu_ms:.0f} ms "
                f"(1 thread), GPU {rec['small']['gpu_ms_host_grid']:.2f} ms with upload and "
                f"download, equal: {rec['small']['equal_to_oracle']}")
        recs["mvdc" if mv else "sparse"] = rec
        del sp
    return recs


def bench_sparse_config(args):
    """The sparse leg alone on the csg8 grid of --size (the default run's sparse object)."""
    import torch
    from contouring_hip import synth
    g = synth.csg8(args.size, seed=args.seed)
    phi = torch.from_numpy(g.values).to("cuda:0")
    recs = bench_sparse(g, phi, args)
    print(json.dumps({"metric": "sparse brick extraction ms (csg8, brick 8)", "config": "sparse",
                      "unit": "ms", "value": recs["mvdc"]["gpu_ms"], "higher_is_better": False,
                      "results": recs}), flush=True)


def bench_m2s_config(args):
    """The mesh-to-SDF leg alone: the csg8 grid of --size extracted (plain FP64), its mesh sampled
    back on the same grid (the default run's mesh_to_sdf object)."""
    import torch
    from contouring_hip import HipDualContouringBackend, _lib, synth
    g = synth.csg8(args.size, seed=args.seed)
    phi = torch.from_numpy(g.values).to("cuda:0")
    be = HipDualContouringBackend("plain-fp64", device=0)
    cg = be._cgrid(g, phi.data_ptr())
    st = _lib.DcHipStats()
    rec = bench_mesh_to_sdf(be, be._lib, cg, _lib.DC_HIP_PLAIN_FP64, st, args)
    print(json.dumps({"metric": "mesh to SDF ms (csg8 DC mesh on its grid)", "config": "m2s",
                      "unit": "ms", "value": rec["gpu_ms"], "higher_is_better": False,
                      "results": rec}), flush=True)


def host_cpu():
    """Host core count and CPU model (for the CPU baseline)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count() or 1
    return {"nproc": nproc, "model": model}


def time_host_input(be, g, args):
    """The reference's extract path: host grid in (upload), host mesh out (download), as
    dc_hip_extract + dc_hip_copy_result into host arrays (CudaDualContouringBackend::extract,
    dc_cuda.cu:614-649, 797-803). Pageable numpy input, as the reference's std::vector."""
    import ctypes

    from contouring_hip import _lib
    lib = be._lib
    vals = np.ascontiguousarray(g.values, np.float32)
    cg = be._cgrid(g, vals.ctypes.data)
    be.set_stream(None)
    be.set_option("timing", 1)
    st = _lib.DcHipStats()
    rows = []
    for q in range(max(2, args.host_steps) + 1):
        t0 = time.perf_counter()
        be._check(lib.dc_hip_extract(be._ctx, ctypes.byref(cg), _lib.DC_HIP_PLAIN_FP64,
                                     ctypes.byref(st)))
        t1 = time.perf_counter()
        V = np.empty(3 * st.vertex_count, np.float32)
        F = np.empty(4 * st.face_count, np.int32)
        be._check(lib.dc_hip_copy_result(be._ctx, V.ctypes.data_as(ctypes.c_void_p), None,
                                         F.ctypes.data_as(ctypes.c_void_p)))
        t2 = time.perf_counter()
        if q > 0:
            rows.append(((t2 - t0) * 1e3, st.upload_ms, (t2 - t1) * 1e3))
    tot, up, down = (float(np.median([r[i] for r in rows])) for i in range(3))
    log(f"[bench] host input: {tot:.2f} ms per extraction (upload {up:.2f} ms, mesh download "
        f"{down:.2f} ms)")
    return {"ms": tot, "upload_ms": up, "download_ms": down, "steps": len(rows),
            "note": "host grid in (pageable), host mesh out; never the headline value"}


def _timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def _device_grid(g):
    import torch
    from contouring_hip import DenseSdfGridDevice
    phi = torch.from_numpy(g.values).to("cuda:0")
    return phi, DenseSdfGridDevice(g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz,
                                   d_values=phi)


def _fwd_bwd_ms(n, steps, warmup, options=None):
    """Plain forward ms and FP32+fallback forward+adjoint ms on the csg8 grid of size n: medians
    over interleaved rounds of 5 timed steps each (at least 10 rounds)."""
    import torch
    from contouring_hip import HipDualContouringBackend, _lib, synth
    g = synth.csg8(n, seed=1234)
    phi, _ = _device_grid(g)
    be = HipDualContouringBackend("plain-fp64", device=0)
    # a created stream (default priority): the adjoint then returns once enqueued, like any
    # stream-ordered library call; the plain forward is timed on the same stream
    stream = torch.cuda.Stream(device=0)
    be.set_stream(stream.cuda_stream)
    be.set_option("timing", 0)  # replayed launch sequences
    torch.cuda.synchronize()  # the grid was uploaded on the default stream
    for key, value in (options or {}).items():
        be.set_option(key, value)
    cg = be._cgrid(g, phi.data_ptr())
    st = _lib.DcHipStats()
    lib = be._lib

    def fwd(prec):
        be._check(lib.dc_hip_extract_device(be._ctx, ctypes.byref(cg), prec, ctypes.byref(st)))

    fwd(_lib.DC_HIP_FP32_FALLBACK)
    nv = int(st.vertex_count)
    gen = torch.Generator(device="cuda:0")
    gen.manual_seed(0)
    gv = torch.randn((nv, 3), dtype=torch.float32, device="cuda:0", generator=gen)
    gphi = torch.empty_like(phi)
    torch.cuda.synchronize()  # made on the default stream, used on `stream`

    def fb():
        fwd(_lib.DC_HIP_FP32_FALLBACK)
        be._check(lib.dc_hip_backward(be._ctx, ctypes.c_void_p(gv.data_ptr()),
                                      ctypes.c_void_p(gphi.data_ptr()), None))

    # interleaved rounds, median per configuration: host and clock drift hit both alike
    plain_r, both_r = [], []
    for _ in range(max(steps, 10)):
        plain_r.append(_timed(lambda: fwd(_lib.DC_HIP_PLAIN_FP64), 5, 2))
        both_r.append(_timed(fb, 5, 2))
    plain_r.sort()
    both_r.sort()
    return plain_r[len(plain_r) // 2], both_r[len(both_r) // 2], nv


def bench_c5(args):
    """Forward + adjoint on 8-box CSG grids at 128^3, 256^3 and 512^3 (loss = sum w_i v_i,
    w ~ N(0,1))."""
    rows = {}
    for n in (128, 256, 512):
        plain, both, nv = _fwd_bwd_ms(n, args.steps, args.warmup)
        rows[n] = {"plain_fwd_ms": plain, "fp32_fwd_bwd_ms": both, "ratio": both / plain,
                   "vertices": nv}
        log(f"[c5] {n}^3: plain {plain:.3f} ms, fp32 fwd+bwd {both:.3f} ms ({both / plain:.2f}x)")
    print(json.dumps({"metric": "fwd+adjoint ms over plain fwd ms (csg8)", "config": "c5",
                      "unit": "ms", "results": rows, "higher_is_better": False}), flush=True)


def bench_c1(args):
    """64^3 spheres: the unit-cube sphere (r 0.35) and dc_cli --generate sphere -g 64 ([-1,1]^3,
    r 0.5). Plain FP64 ms per extraction (replayed launch sequence), active cells/s, and parity
    with the oracle: quads and FP32 vertices exact, FP64 vertices within 1e-6 h."""
    import numpy as np
    import torch
    from contouring_hip import HipDualContouringBackend, _lib, synth
    import oracle
    rows = {}
    for g in (synth.sphere_unit(64), synth.cli_generate("sphere", 64)):
        phi, _ = _device_grid(g)
        be = HipDualContouringBackend("plain-fp64", device=0)
        be.set_stream(torch.cuda.current_stream(0).cuda_stream)
        be.set_option("timing", 0)
        cg = be._cgrid(g, phi.data_ptr())
        st = _lib.DcHipStats()
        lib = be._lib
        ms = _timed(lambda: be._check(lib.dc_hip_extract_device(be._ctx, ctypes.byref(cg),
                                                                _lib.DC_HIP_PLAIN_FP64,
                                                                ctypes.byref(st))),
                    args.steps, max(args.warmup, 3))
        nv, nf = int(st.vertex_count), int(st.face_count)
        V32 = np.empty((nv, 3), np.float32)
        V64 = np.empty((nv, 3), np.float64)
        F = np.empty((nf, 4), np.int32)
        be._check(lib.dc_hip_copy_result(be._ctx, V32.ctypes.data_as(ctypes.c_void_p),
                                         V64.ctypes.data_as(ctypes.c_void_p),
                                         F.ctypes.data_as(ctypes.c_void_p)))
        ref = oracle.run_grid(g)
        h = min(g.vx, g.vy, g.vz)
        ok = (nv == ref.n_vertices and nf == ref.n_faces and np.array_equal(F, ref.F)
              and np.array_equal(V32, ref.V32) and float(np.abs(V64 - ref.V64).max()) <= 1e-6 * h)
        rows[g.name] = {"ms": ms, "vertices": nv, "faces": nf,
                        "active_cells_per_s": nv / (ms * 1e-3), "parity": bool(ok),
                        "launch": "hipGraph replay" if st.graph_mode == 2 else "eager"}
        log(f"[c1] {g.name}: {ms:.3f} ms, {nv} vertices, {nv / (ms * 1e-3) / 1e6:.1f} M active "
            f"cells/s, parity {ok}")
    print(json.dumps({"metric": "64^3 sphere DC ms (plain FP64), parity with the oracle",
                      "config": "c1", "unit": "ms", "results": rows, "higher_is_better": False}),
          flush=True)


def bench_c3(args):
    """1024^3 narrow band (64-sphere smooth-min, clamped to +-4h): plain vs FP32+fallback time and
    the FP32 vertex error against the plain FP64 vertices."""
    import torch
    from contouring_hip import HipDualContouringBackend, _lib, synth
    n = 1024
    t = time.time()
    g = synth.blend64(n)
    log(f"[c3] generated blend64 {n}^3 in {time.time() - t:.1f}s")
    phi, _ = _device_grid(g)
    del g.values
    be = HipDualContouringBackend("plain-fp64", device=0)
    be.set_stream(torch.cuda.current_stream(0).cuda_stream)
    cg = be._cgrid(g, phi.data_ptr())
    st = _lib.DcHipStats()
    lib = be._lib
    res = {}
    for name, prec in (("plain-fp64", _lib.DC_HIP_PLAIN_FP64),
                       ("fp32-fallback", _lib.DC_HIP_FP32_FALLBACK)):
        ms = _timed(lambda: be._check(lib.dc_hip_extract_device(be._ctx, ctypes.byref(cg), prec,
                                                                ctypes.byref(st))),
                    max(3, args.steps // 4), 1)
        nv = int(st.vertex_count)
        res[name] = {"ms": ms, "vertices": nv, "faces": int(st.face_count),
                     "fallback_cells": int(st.qef_fallback_count),
                     "active_fraction": nv / float((n - 1) ** 3)}
        if name == "plain-fp64":
            v64 = torch.empty((nv, 3), dtype=torch.float64, device="cuda:0")
            be._check(lib.dc_hip_copy_result(be._ctx, None, ctypes.c_void_p(v64.data_ptr()),
                                             None))
        else:
            v32 = torch.empty((nv, 3), dtype=torch.float32, device="cuda:0")
            be._check(lib.dc_hip_copy_result(be._ctx, ctypes.c_void_p(v32.data_ptr()), None,
                                             None))
        log(f"[c3] {name}: {ms:.3f} ms, {nv} vertices")
    err = (v32.double() - v64).abs() / g.h
    res["fp32_vertex_err_h"] = {"max": float(err.max()), "rms": float(err.pow(2).mean().sqrt())}
    log(f"[c3] fp32 vertex error: max {res['fp32_vertex_err_h']['max']:.2e} h, "
        f"rms {res['fp32_vertex_err_h']['rms']:.2e} h")
    print(json.dumps({"metric": "1024^3 narrow-band DC ms, plain vs FP32+fallback", "config": "c3",
                      "unit": "ms", "results": res, "higher_is_better": False}), flush=True)


def bench_c4(args):
    """256 uneven CSG grids (32^3..256^3): one extract_batch launch sequence vs 256 sequential
    extract_device calls on the same device-resident grids; per-grid spot parity."""
    import torch
    from contouring_hip import HipDualContouringBackend, _lib, synth
    t = time.time()
    grids = synth.batch_grids(256)
    log(f"[c4] generated 256 grids ({sum(g.values.size for g in grids) / 1e9:.2f} G samples) in "
        f"{time.time() - t:.1f}s")
    dev = [_device_grid(g) for g in grids]
    for g in grids:
        g.values = None
    be = HipDualContouringBackend("plain-fp64", device=0)
    be.set_stream(torch.cuda.current_stream(0).cuda_stream)
    be.set_option("timing", 0)  # the batch's launch sequence is captured once and replayed
    lib = be._lib
    arr = (_lib.DcHipGrid * len(grids))(*[be._cgrid(g, d[0].data_ptr())
                                           for g, d in zip(grids, dev)])
    voff = (ctypes.c_int64 * (len(grids) + 1))()
    foff = (ctypes.c_int64 * (len(grids) + 1))()
    st = _lib.DcHipStats()
    prec = _lib.DC_HIP_PLAIN_FP64

    def batch():
        be._check(lib.dc_hip_extract_batch(be._ctx, arr, len(grids), prec, voff, foff,
                                           ctypes.byref(st)))

    steps = max(3, args.steps // 4)
    batch_ms = _timed(batch, steps, 2)
    tot_v, tot_f = int(st.vertex_count), int(st.face_count)
    spot = [0, 17, 101, 255]
    batch_v = torch.empty((tot_v, 3), dtype=torch.float64, device="cuda:0")
    batch_f = torch.empty((tot_f, 4), dtype=torch.int32, device="cuda:0")
    be._check(lib.dc_hip_copy_result(be._ctx, None, ctypes.c_void_p(batch_v.data_ptr()),
                                     ctypes.c_void_p(batch_f.data_ptr())))
    offs = [(voff[q], voff[q + 1], foff[q], foff[q + 1]) for q in range(len(grids))]

    def sequential():
        for q in range(len(grids)):
            be._check(lib.dc_hip_extract_device(be._ctx, ctypes.byref(arr[q]), prec,
                                                ctypes.byref(st)))

    seq_ms = _timed(sequential, steps, 1)
    ok = True
    for q in spot:
        be._check(lib.dc_hip_extract_device(be._ctx, ctypes.byref(arr[q]), prec,
                                            ctypes.byref(st)))
        nv, nf = int(st.vertex_count), int(st.face_count)
        v = torch.empty((nv, 3), dtype=torch.float64, device="cuda:0")
        f = torch.empty((nf, 4), dtype=torch.int32, device="cuda:0")
        be._check(lib.dc_hip_copy_result(be._ctx, None, ctypes.c_void_p(v.data_ptr()),
                                         ctypes.c_void_p(f.data_ptr())))
        a, b, c, d = offs[q]
        ok &= (b - a == nv and d - c == nf and torch.equal(batch_v[a:b], v)
               and torch.equal(batch_f[c:d], f))
    log(f"[c4] batch {batch_ms:.3f} ms vs sequential {seq_ms:.3f} ms "
        f"({seq_ms / batch_ms:.2f}x), {tot_v} vertices, spot parity {ok}")
    print(json.dumps({"metric": "256 uneven grids: one batch vs 256 sequential extractions, ms",
                      "config": "c4", "unit": "ms", "batch_ms": batch_ms, "sequential_ms": seq_ms,
                      "speedup": seq_ms / batch_ms, "vertices": tot_v, "faces": tot_f,
                      "spot_parity": bool(ok), "higher_is_better": False}), flush=True)


if __name__ == "__main__":
    main()
