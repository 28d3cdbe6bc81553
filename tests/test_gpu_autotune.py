"""The marking pass's tile-shape sweep (dc_hip_autotune through contouring_hip.autotune_mark).

The tile shape changes only the launch geometry: the tuned context must give the same mesh as the
default one, bit for bit. The sweep times all 12 candidates, keeps the fastest on the context and
restores the caller's timing option.
"""
import numpy as np
import pytest

from contouring_hip import DenseSdfGridDevice, HipDualContouringBackend, autotune_mark, synth

pytestmark = pytest.mark.gpu


def test_autotune_keeps_results_and_picks_fastest(tmp_path):
    import torch
    g = synth.csg8(96, seed=21)
    phi = torch.from_numpy(np.ascontiguousarray(g.values, np.float32)).cuda()
    dev = DenseSdfGridDevice(g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz, d_values=phi)
    ref = HipDualContouringBackend("plain-fp64")
    (V0, F0), _ = ref.extract_tensor(phi.reshape(g.nz, g.ny, g.nx), (g.vx, g.vy, g.vz),
                                     (g.ox, g.oy, g.oz))
    be = HipDualContouringBackend("plain-fp64")
    be.set_option("timing", 0)
    # a fresh cache file: the first call sweeps and stores its choice
    best, table = autotune_mark(be, dev, repeats=2, cache_path=str(tmp_path / "tune.json"))
    print(f"[autotune] best {best}", flush=True)
    assert len(table) == 12 and all(row["mark_ms"] > 0 for row in table)
    assert best["mark_ms"] == min(row["mark_ms"] for row in table)
    (V1, F1), st = be.extract_tensor(phi.reshape(g.nz, g.ny, g.nx), (g.vx, g.vy, g.vz),
                                     (g.ox, g.oy, g.oz))
    assert torch.equal(V0, V1) and torch.equal(F0, F1)
    assert st.marking_ms == 0  # timing stayed off
    # the disk cache gives the same choice without a sweep
    best2, table2 = autotune_mark(be, dev, cache_path=str(tmp_path / "tune.json"), use_cache=True)
    assert table2 == [] and (best2["rows"], best2["waves"]) == (best["rows"], best["waves"])


def test_autotune_launch_keeps_results_and_picks_fastest(tmp_path):
    """The per-cell launch shapes (solve blocks, gather rows per wave) are picked by whole timed
    steps: the mesh stays bit-identical, the gradient within the gather's 1e-6 (relative) order
    difference, the fastest candidate is kept, the timing option is restored and the disk cache
    returns the same choice without a sweep."""
    import torch
    g = synth.csg8(96, seed=21)
    phi = torch.from_numpy(np.ascontiguousarray(g.values, np.float32)).cuda()
    phi3 = phi.reshape(g.nz, g.ny, g.nx)
    dev = DenseSdfGridDevice(g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz, d_values=phi)
    sp, org = (g.vx, g.vy, g.vz), (g.ox, g.oy, g.oz)
    ref = HipDualContouringBackend("fp32-fallback")
    (V0, F0), _ = ref.extract_tensor(phi3, sp, org)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    w = torch.randn(V0.shape, device="cuda", generator=gen)
    g0 = torch.full_like(phi3, float("nan"))
    ref.backward(w, g0)
    be = HipDualContouringBackend("fp32-fallback")
    cache = str(tmp_path / "tune.json")
    best, table = be.autotune_launch(dev, repeats=2, adjoint=True, cache_path=cache)
    print(f"[autotune] launch best {best}", flush=True)
    assert len(table) == 12 and all(row["step_ms"] > 0 for row in table)
    assert best["step_ms"] == min(row["step_ms"] for row in table)
    (V1, F1), st = be.extract_tensor(phi3, sp, org)
    assert torch.equal(V0, V1) and torch.equal(F0, F1)
    assert st.marking_ms > 0  # per-pass timing (the default) is on again
    g1 = torch.full_like(phi3, float("nan"))
    be.backward(w, g1)
    assert torch.isfinite(g1).all()
    rel = float((g1.double() - g0.double()).abs().max() / g0.double().abs().max())
    assert rel <= 1e-6, rel
    best2, table2 = be.autotune_launch(dev, adjoint=True, cache_path=cache)
    assert table2 == [] and best2 == best
    # a device pointer in place of the tensor, forward shapes only
    dev_ptr = DenseSdfGridDevice(g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz,
                                 d_values=phi.data_ptr())
    best3, table3 = be.autotune_launch(dev_ptr, repeats=1, cache_path=str(tmp_path / "t3.json"))
    assert len(table3) == 4 and best3["gather_rows"] == 1
