"""Captured forward launch sequences (hipGraph) against eager runs.

With per-pass timing off, a context captures a plan the second time in a row it runs and replays
it while the plan is unchanged. Every replay must give the eager result bit for bit, with other GPU
work and host synchronisation between the replays, on the null stream and on a created stream; the
adjoint after a replayed forward must equal the adjoint after an eager forward.
"""
import numpy as np
import pytest

from contouring_hip import DenseSdfGridDevice, HipDualContouringBackend, synth

pytestmark = pytest.mark.gpu


def _phi(g):
    import torch
    return torch.from_numpy(np.ascontiguousarray(g.values, np.float32)).cuda().reshape(
        g.nz, g.ny, g.nx)


def _graph_backend(precision):
    be = HipDualContouringBackend(precision)
    be.set_option("timing", 0)  # graphs are used only without per-pass events
    return be


def _grids():
    g1, g2 = synth.csg8(64, seed=1234), synth.sphere_unit(48)
    return [(g, _phi(g)) for g in (g1, g2)]


def _run(be, g, phi):
    f64 = be.precision == "plain-fp64"
    out, st = be.extract_tensor(phi, (g.vx, g.vy, g.vz), (g.ox, g.oy, g.oz), with_f64=f64)
    return (out if f64 else out + (None,)), st


def _check_sequence(precision):
    import torch
    items = _grids()
    eager = HipDualContouringBackend(precision)
    eager.set_option("graph", 0)
    refs = [_run(eager, g, phi)[0] for g, phi in items]
    be = _graph_backend(precision)
    # eager, capture, replay, replay, other plan (eager), replay of the first plan's graph, ...
    modes = [0, 1, 2, 2, 0, 2, 2, 2]
    for step, idx in enumerate([0, 0, 0, 0, 1, 0, 0, 0]):
        g, phi = items[idx]
        print(f"[graph] {precision} step {step} grid {g.name}", flush=True)
        (V, F, V64), st = _run(be, g, phi)
        rV, rF, rV64 = refs[idx]
        assert torch.equal(V, rV) and torch.equal(F, rF), (step, g.name)
        if precision == "plain-fp64":
            assert torch.equal(V64, rV64), (step, g.name)
        assert st.vertex_count == rV.shape[0] and st.face_count == rF.shape[0]
        assert st.graph_mode == modes[step], (step, st.graph_mode)
        # unrelated GPU work and a host synchronisation between replays
        junk = torch.randn(1 << 20, device="cuda") * 3.0
        assert torch.isfinite(junk.sum()).item()
        torch.cuda.synchronize()


@pytest.mark.parametrize("precision", ["plain-fp64", "fp32-fallback"])
def test_graph_replays_match_eager_null_stream(precision):
    _check_sequence(precision)


def test_graph_replays_match_eager_created_stream():
    import torch
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        _check_sequence("plain-fp64")
    torch.cuda.synchronize()


def test_adjoint_after_replay_matches_eager():
    import torch
    g, phi = _grids()[0]
    eager = HipDualContouringBackend("fp32-fallback")
    eager.set_option("graph", 0)
    be = _graph_backend("fp32-fallback")
    (V, F), _ = eager.extract_tensor(phi, (g.vx, g.vy, g.vz), (g.ox, g.oy, g.oz))
    gen = torch.Generator(device="cuda")
    gen.manual_seed(7)
    w = torch.randn(V.shape, device="cuda", generator=gen)
    want = torch.empty_like(phi)
    eager.backward(w, want)
    got = torch.empty_like(phi)
    for rep in range(4):  # forward and adjoint each: eager, capture, replay, replay
        print(f"[graph] adjoint rep {rep}", flush=True)
        (V2, F2), st = be.extract_tensor(phi, (g.vx, g.vy, g.vz), (g.ox, g.oy, g.oz))
        assert st.graph_mode == [0, 1, 2, 2][rep]
        assert torch.equal(V2, V) and torch.equal(F2, F)
        got.fill_(float("nan"))
        bst = be.backward(w, got)
        assert bst.graph_mode == [0, 1, 2, 2][rep]
        assert torch.equal(got, want), rep
        torch.cuda.synchronize()


def test_batch_replays_match_eager():
    import torch
    items = _grids() + [(synth.csg8(40, seed=9), None)]
    items[2] = (items[2][0], _phi(items[2][0]))
    dev = [DenseSdfGridDevice(g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz,
                              d_values=phi.reshape(-1)) for g, phi in items]
    eager = HipDualContouringBackend("plain-fp64")
    eager.set_option("graph", 0)
    ref, _ = eager.extract_batch(dev)
    be = _graph_backend("plain-fp64")
    for rep in range(4):
        print(f"[graph] batch rep {rep}", flush=True)
        pairs, _ = be.extract_batch(dev)
        for (V, F), (rV, rF) in zip(pairs, ref):
            assert torch.equal(V, rV) and torch.equal(F, rF), rep
        torch.cuda.synchronize()


def test_adjoint_enqueued_on_caller_stream():
    """With timing off on a caller's stream the adjoint returns once enqueued; work queued after
    it on that stream (here: a copy and a comparison) sees the complete gradient, on every
    launch form (eager, capture, replay)."""
    import torch
    g, phi = _grids()[0]
    eager = HipDualContouringBackend("fp32-fallback")
    eager.set_option("graph", 0)
    (V, F), _ = eager.extract_tensor(phi, (g.vx, g.vy, g.vz), (g.ox, g.oy, g.oz))
    gen = torch.Generator(device="cuda")
    gen.manual_seed(3)
    w = torch.randn(V.shape, device="cuda", generator=gen)
    want = torch.empty_like(phi)
    eager.backward(w, want)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    be = _graph_backend("fp32-fallback")
    with torch.cuda.stream(s):
        got = torch.empty_like(phi)
        for rep in range(4):
            print(f"[graph] enqueued adjoint rep {rep}", flush=True)
            be.extract_tensor(phi, (g.vx, g.vy, g.vz), (g.ox, g.oy, g.oz))
            got.fill_(float("nan"))
            bst = be.backward(w, got)
            copy = got.clone()  # queued on s right behind the adjoint
            assert bst.graph_mode == [0, 1, 2, 2][rep]
            assert torch.equal(copy, want), rep
    torch.cuda.synchronize()


def test_null_stream_inputs_are_ordered():
    """A call from the null stream runs after the work already queued there: a grid produced by
    a long null-stream kernel sequence just before the call extracts like the finished grid."""
    import torch
    g, phi = _grids()[0]
    be = HipDualContouringBackend("plain-fp64")
    (V, F), _ = be.extract_tensor(phi, (g.vx, g.vy, g.vz), (g.ox, g.oy, g.oz))
    torch.cuda.synchronize()
    for rep in range(3):
        src = phi.clone()
        torch.cuda.synchronize()
        tmp = torch.full_like(phi, float("nan"))
        big = torch.randn(1 << 24, device="cuda")
        for _ in range(8):  # keep the null stream busy before the grid is written
            big = big * 1.0001
        tmp.copy_(src)  # the grid, written last on the null stream
        (V2, F2), _ = be.extract_tensor(tmp, (g.vx, g.vy, g.vz), (g.ox, g.oy, g.oz))
        assert torch.equal(V2, V) and torch.equal(F2, F), rep
