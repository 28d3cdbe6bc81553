"""FP32+FP64-fallback mode and the adjoint on the GPU, against the FP64 oracle.

Tolerances: FP32 mode vertices within 1e-4*h of the FP64 oracle (the stated target); topology
exact: active set, vertex ids and every quad with its orientation equal to the oracle's (quads
whose orientation test is within the FP32 vertices' error of a tie are oriented again from FP64
vertices, k_reorient).
Adjoint: the plain-FP64 adjoint within 1e-5 relative (max-norm) of the oracle's FP64 analytic
gradient, the FP32 adjoint within 2e-3; the gradient is checked at fixed topology.
"""
import numpy as np
import pytest

import oracle
from contouring_hip import DenseSdfGrid, HipDualContouringBackend, synth
from grids import ragged_sphere

pytestmark = pytest.mark.gpu

_be32 = None


def be32():
    global _be32
    if _be32 is None:
        _be32 = HipDualContouringBackend("fp32-fallback")
    return _be32


def canon(F):
    F = F.reshape(-1, 4).copy()
    lo = np.minimum(F[:, 1], F[:, 3])
    hi = np.maximum(F[:, 1], F[:, 3])
    F[:, 1], F[:, 3] = lo, hi
    return F


@pytest.mark.parametrize("make", [lambda: synth.csg8(64, seed=1234),
                                  lambda: synth.csg8(128, seed=42),
                                  lambda: synth.sphere_unit(64),
                                  lambda: synth.blend64(96, seed=5)])
def test_fp32_fallback_mode_error(make):
    g = make()
    print(f"[fp32] {g.name}", flush=True)
    ref = oracle.run_grid(g)
    grid = DenseSdfGrid(g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz, g.values)
    mesh, st = be32().extract(grid)
    assert st.vertex_count == ref.n_vertices
    assert st.face_count == ref.n_faces
    F = mesh.faces.reshape(-1, 4)
    np.testing.assert_array_equal(canon(F), canon(ref.F))
    flips = int(np.sum(np.any(F != ref.F, axis=1)))
    np.testing.assert_array_equal(F, ref.F)
    err = np.abs(mesh.vertices.reshape(-1, 3).astype(np.float64) - ref.V64) / g.h
    print(f"[fp32] {g.name}: max {err.max():.3e} h, rms {np.sqrt((err ** 2).mean()):.3e} h, "
          f"fallback {st.qef_fallback_count}/{st.vertex_count}, orientation flips {flips}",
          flush=True)
    assert err.max() <= 1e-4


@pytest.mark.parametrize("precision,tol", [("plain-fp64", 1e-5), ("fp32-fallback", 2e-3)])
def test_adjoint_matches_oracle_gradient(precision, tol):
    import torch
    g = synth.csg8(32, seed=1234)
    phi = torch.from_numpy(np.asarray(g.values, np.float32)).reshape(g.nz, g.ny, g.nx).cuda()
    be = HipDualContouringBackend(precision)
    (V, F), _ = be.extract_tensor(phi, (g.vx, g.vy, g.vz), (g.ox, g.oy, g.oz))
    ref = oracle.run_grid(g)
    assert V.shape[0] == ref.n_vertices
    rng = np.random.default_rng(0)
    w = rng.normal(size=(ref.n_vertices, 3))
    gphi = torch.empty_like(phi)
    be.backward(torch.from_numpy(w.astype(np.float32)).cuda(), gphi)
    got = gphi.cpu().numpy().ravel().astype(np.float64)
    want = ref.backward(w)
    rel = np.abs(got - want).max() / np.abs(want).max()
    print(f"[adjoint] {precision}: max rel err {rel:.3e}", flush=True)
    assert rel <= tol
    # samples outside every active cell's stencil get exactly zero
    far = np.ones(want.shape, bool)
    nx, ny, nz = g.nx, g.ny, g.nz
    cx, cy = nx - 1, ny - 1
    for c in ref.active:
        i, j, k = c % cx, (c // cx) % cy, c // (cx * cy)
        for dz in range(-1, 3):
            for dy in range(-1, 3):
                for dx in range(-1, 3):
                    x, y, z = i + dx, j + dy, k + dz
                    if 0 <= x < nx and 0 <= y < ny and 0 <= z < nz:
                        far[x + nx * (y + ny * z)] = False
    assert np.all(got[far] == 0)


def test_adjoint_deterministic_and_autograd():
    import torch
    from contouring_hip.autograd import dual_contour
    g = synth.csg8(48, seed=3)
    phi = torch.from_numpy(np.asarray(g.values, np.float32)).reshape(g.nz, g.ny, g.nx).cuda()
    phi.requires_grad_(True)
    be = HipDualContouringBackend("fp32-fallback")
    w = torch.randn(1, device="cuda")  # placeholder to seed the CUDA RNG state
    grads = []
    for _ in range(2):
        V, F = dual_contour(phi, (g.vx, g.vy, g.vz), (g.ox, g.oy, g.oz), backend=be)
        torch.manual_seed(0)
        wv = torch.randn(V.shape, device="cuda")
        (V * wv).sum().backward()
        grads.append(phi.grad.clone())
        phi.grad = None
    assert torch.equal(grads[0], grads[1])
    assert torch.isfinite(grads[0]).all()
    del w


@pytest.mark.parametrize("kept", [0, 5])
def test_adjoint_fallback_state_capacity(kept):
    """FP32 mode keeps the FP64 fallback solves' states for up to `fallback_state` FP64 cells;
    with more, the adjoint solves them all again. Either way the gradient is the oracle's to the
    FP32 tolerance, the default capacity's (states used) to 1e-4 (relative, max-norm), and the
    same on a second context."""
    import torch
    g = synth.csg8(40, seed=77)
    phi = torch.from_numpy(np.asarray(g.values, np.float32)).reshape(g.nz, g.ny, g.nx).cuda()
    rng = np.random.default_rng(3)
    grads = []
    for opt in (-1, kept, kept):
        be = HipDualContouringBackend("fp32-fallback")
        be.set_option("fallback_state", opt)
        (V, F), st = be.extract_tensor(phi, (g.vx, g.vy, g.vz), (g.ox, g.oy, g.oz))
        assert st.qef_fallback_count > kept  # some FP64 cells are solved again
        if not grads:
            w = torch.from_numpy(rng.normal(size=(V.shape[0], 3)).astype(np.float32)).cuda()
        gphi = torch.empty_like(phi)
        be.backward(w, gphi)
        grads.append(gphi.cpu().numpy().ravel().astype(np.float64))
    print(f"[adjoint] fallback_state {kept}: {st.qef_fallback_count} FP64 cells", flush=True)
    ref = oracle.run_grid(g)
    want = ref.backward(w.cpu().numpy().astype(np.float64))
    scale = np.abs(want).max()
    assert np.abs(grads[1] - want).max() / scale <= 2e-3
    assert np.abs(grads[1] - grads[0]).max() / scale <= 1e-4
    assert np.array_equal(grads[1], grads[2])  # reproducible across contexts


@pytest.mark.parametrize("precision", ["plain-fp64", "fp32-fallback"])
def test_adjoint_lanes_agree(precision):
    """The per-cell adjoint kernels at one thread per cell and at 4 and 16 lanes per cell (edges
    split over the lanes) sum the same terms in other orders: equal to 1e-5 (relative,
    max-norm), each reproducible."""
    import torch
    g = synth.csg8(36, seed=21)
    phi = torch.from_numpy(np.asarray(g.values, np.float32)).reshape(g.nz, g.ny, g.nx).cuda()
    grads = {}
    for lanes in (1, 4, 16, 16):
        be = HipDualContouringBackend(precision)
        be.set_option("adjoint_lanes", lanes)
        be.set_option("adjoint_face_lanes", lanes)
        be.set_option("adjoint_fallback_lanes", 1 if lanes == 1 else 16)
        (V, F), _ = be.extract_tensor(phi, (g.vx, g.vy, g.vz), (g.ox, g.oy, g.oz))
        gen = torch.Generator(device="cuda")
        gen.manual_seed(5)
        w = torch.randn(V.shape, device="cuda", generator=gen)
        gphi = torch.empty_like(phi)
        be.backward(w, gphi)
        got = gphi.cpu().numpy().ravel().astype(np.float64)
        if lanes in grads:
            assert np.array_equal(got, grads[lanes])
        grads[lanes] = got
    scale = np.abs(grads[1]).max()
    for lanes in (4, 16):
        rel = np.abs(grads[lanes] - grads[1]).max() / scale
        print(f"[adjoint] {precision}: lanes 1 vs {lanes} rel {rel:.2e}", flush=True)
        assert rel <= 1e-5


def test_gather_blocks_bit_identical():
    """The gather with a bounded grid looping over its row groups (option gather_blocks) gives the
    one-wave-per-group gradient bit for bit: every row is summed by one wave in the same order."""
    import torch
    g = synth.csg8(48, seed=5)
    phi = torch.from_numpy(np.asarray(g.values, np.float32)).reshape(g.nz, g.ny, g.nx).cuda()
    got = []
    for blocks in (0, 1, 7, 64):
        be = HipDualContouringBackend("fp32-fallback")
        be.set_option("gather_blocks", blocks)
        (V, F), _ = be.extract_tensor(phi, (g.vx, g.vy, g.vz), (g.ox, g.oy, g.oz))
        w = torch.ones(V.shape, device="cuda")
        gphi = torch.full_like(phi, float("nan"))
        be.backward(w, gphi)
        r = gphi.cpu().numpy()
        assert np.isfinite(r).all(), blocks
        got.append(r)
    for q in range(1, len(got)):
        assert np.array_equal(got[q], got[0])


def test_gather_rows_agree():
    """The gradient gather at 1, 2, 4 and 8 sample rows per wave adds each sample's (at most 8)
    corner terms in orders that may differ: equal to 1e-6 (relative, max-norm) and each
    reproducible. The plain-mode gradient matches the oracle's whichever is set: to 1e-5 on the
    isotropic grid, and to 1e-4 on the 1100 x 14 x 13 grid. Its spacings differ 85-fold, and
    the FP32 corner terms, large and partly cancelling, leave ~5e-5 (the error is spread over x,
    not at the 1024-sample step boundary this grid crosses)."""
    import torch
    for g, tol in ((synth.csg8(40, seed=13), 1e-5), (ragged_sphere(1100, 14, 13), 1e-4)):
        phi = torch.from_numpy(np.asarray(g.values, np.float32)).reshape(g.nz, g.ny, g.nx).cuda()
        got = {}
        for rows in (1, 2, 4, 8, 8):
            be = HipDualContouringBackend("plain-fp64")
            be.set_option("gather_rows", rows)
            (V, F), _ = be.extract_tensor(phi, (g.vx, g.vy, g.vz), (g.ox, g.oy, g.oz))
            gen = torch.Generator(device="cuda")
            gen.manual_seed(11)
            w = torch.randn(V.shape, device="cuda", generator=gen)
            gphi = torch.full_like(phi, float("nan"))
            be.backward(w, gphi)
            r = gphi.cpu().numpy().ravel().astype(np.float64)
            assert np.isfinite(r).all(), (g.name, rows)
            if rows in got:
                assert np.array_equal(r, got[rows]), (g.name, rows)
            got[rows] = r
        scale = np.abs(got[1]).max()
        for rows in (2, 4, 8):
            rel = np.abs(got[rows] - got[1]).max() / scale
            print(f"[gather] {g.name}: rows {rows} vs 1 rel {rel:.2e}", flush=True)
            assert rel <= 1e-6, (g.name, rows)
        ref = oracle.run_grid(g)
        want = ref.backward(w.cpu().numpy().astype(np.float64))
        assert np.abs(got[8] - want).max() / np.abs(want).max() <= tol, g.name
