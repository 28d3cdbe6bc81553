"""The Python boundary's tensor handling on the GPU.

- dL/dphi of DualContouringFunction is laid out like phi's logical shape whatever phi's strides
  are (a permuted phi gets the same gradient as its contiguous copy, and the oracle's);
- two forwards on one backend before the first forward's backward rebuild that forward's topology
  (the generation-mismatch path) and give the oracle's gradient;
- tensors of the wrong dtype, device layout or size are rejected with ValueError before anything
  is launched (their pointers would otherwise be indexed by the grid's sizes).
"""
import numpy as np
import pytest

import oracle
from contouring_hip import DenseSdfGridDevice, HipDualContouringBackend, synth

pytestmark = pytest.mark.gpu


def _phi(g):
    import torch
    return torch.from_numpy(np.asarray(g.values, np.float32)).reshape(g.nz, g.ny, g.nx).cuda()


def _weights(nv, seed=0):
    import torch
    rng = np.random.default_rng(seed)
    w = rng.normal(size=(nv, 3))
    return w, torch.from_numpy(w.astype(np.float32)).cuda()


def test_permuted_phi_gradient_matches_contiguous_and_oracle():
    import torch
    from contouring_hip.autograd import dual_contour
    g = synth.csg8(32, seed=1234)
    print("[boundary] permuted phi, csg8 32", flush=True)
    base = _phi(g)
    # the same logical (nz, ny, nx) tensor stored with x slowest: strides are not contiguous
    perm = base.permute(2, 1, 0).contiguous().permute(2, 1, 0)
    assert not perm.is_contiguous() and torch.equal(perm, base)
    be = HipDualContouringBackend("plain-fp64")
    grads = []
    for phi in (base.clone(), perm.clone()):
        phi.requires_grad_(True)
        V, F = dual_contour(phi, (g.vx, g.vy, g.vz), (g.ox, g.oy, g.oz), backend=be)
        w, wt = _weights(V.shape[0])
        (V * wt).sum().backward()
        grads.append(phi.grad.contiguous().clone())
    assert torch.equal(grads[0], grads[1])
    ref = oracle.run_grid(g)
    want = ref.backward(w)
    got = grads[1].cpu().numpy().ravel().astype(np.float64)
    rel = np.abs(got - want).max() / np.abs(want).max()
    assert rel <= 1e-5, rel


def test_two_forwards_before_backward_reextracts():
    import torch
    from contouring_hip.autograd import dual_contour
    g1, g2 = synth.csg8(32, seed=1234), synth.csg8(28, seed=7)
    print("[boundary] two forwards then backward", flush=True)
    be = HipDualContouringBackend("plain-fp64")
    phi1 = _phi(g1).requires_grad_(True)
    phi2 = _phi(g2).requires_grad_(True)
    V1, _ = dual_contour(phi1, (g1.vx, g1.vy, g1.vz), (g1.ox, g1.oy, g1.oz), backend=be)
    V2, _ = dual_contour(phi2, (g2.vx, g2.vy, g2.vz), (g2.ox, g2.oy, g2.oz), backend=be)
    w, wt = _weights(V1.shape[0], seed=3)
    (V1 * wt).sum().backward()  # the context last held g2's forward
    want = oracle.run_grid(g1).backward(w)
    got = phi1.grad.cpu().numpy().ravel().astype(np.float64)
    assert np.abs(got - want).max() / np.abs(want).max() <= 1e-5
    assert V2.shape[0] == oracle.run_grid(g2).n_vertices


def test_bad_tensors_rejected_before_launch():
    import torch
    g = synth.csg8(24, seed=2)
    print("[boundary] bad tensors", flush=True)
    phi = _phi(g)
    be = HipDualContouringBackend("plain-fp64")
    (V, F), _ = be.extract_tensor(phi, (g.vx, g.vy, g.vz), (g.ox, g.oy, g.oz))
    nv = V.shape[0]
    gv = torch.zeros((nv, 3), device="cuda")
    good = torch.empty_like(phi)
    bad_grads = {
        "short": torch.empty(phi.numel() - 1, device="cuda"),
        "fp16": torch.empty(phi.shape, dtype=torch.float16, device="cuda"),
        "host": torch.empty(phi.shape),
        "strided": torch.empty((phi.numel() * 2,), device="cuda")[::2],
    }
    for name, t in bad_grads.items():
        with pytest.raises(ValueError):
            be.backward(gv, t)
        print(f"[boundary] grad_phi {name}: rejected", flush=True)
    with pytest.raises(ValueError):
        be.backward(torch.zeros((nv + 1, 3), device="cuda"), good)
    with pytest.raises(ValueError):
        be.backward(torch.zeros((nv, 3)), good)  # host tensor
    # device grids: wrong size and dtype
    for vals in (phi.reshape(-1)[:-1], phi.to(torch.float16), phi.reshape(-1)[::2]):
        d = DenseSdfGridDevice(g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz, d_values=vals)
        with pytest.raises(ValueError):
            be.extract_device(d)
        with pytest.raises(ValueError):
            be.extract_batch([d])
    # the context still holds the good forward: a correct backward works and matches the oracle
    w, wt = _weights(nv, seed=5)
    be.backward(wt, good)
    want = oracle.run_grid(g).backward(w)
    got = good.cpu().numpy().ravel().astype(np.float64)
    assert np.abs(got - want).max() / np.abs(want).max() <= 1e-5


def test_gradient_aliasing_the_samples_is_rejected():
    """The adjoint reads the forward's samples again while it writes the gradient: a gradient
    buffer that is the forward's own sample tensor is refused before anything is launched."""
    from contouring_hip import _lib
    g = synth.csg8(24, seed=4)
    print("[boundary] aliasing gradient", flush=True)
    phi = _phi(g)
    be = HipDualContouringBackend("plain-fp64")
    (V, F), _ = be.extract_tensor(phi, (g.vx, g.vy, g.vz), (g.ox, g.oy, g.oz))
    w, wt = _weights(V.shape[0], seed=6)
    with pytest.raises(_lib.DcHipError) as ei:
        be.backward(wt, phi)
    assert ei.value.status == _lib.DC_HIP_ERR_INVALID_ARG
    # a separate gradient still works, and the samples are unchanged
    before = phi.clone()
    good = torch_empty_like(phi)
    be.backward(wt, good)
    assert bool((phi == before).all())
    want = oracle.run_grid(g).backward(w)
    got = good.cpu().numpy().ravel().astype(np.float64)
    assert np.abs(got - want).max() / np.abs(want).max() <= 1e-5


def torch_empty_like(t):
    import torch
    return torch.empty_like(t)
