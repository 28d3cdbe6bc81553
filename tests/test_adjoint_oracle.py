"""Pins the oracle's FP64 analytic adjoint (the reference gradient for the GPU adjoint) against
FP64 central differences of the oracle's forward (step 1e-6), at fixed topology."""
import numpy as np
import pytest

import oracle
from contouring_hip import synth


def _loss_grad_fd(g, w, idx, step=1e-6):
    base = np.asarray(g.values, np.float64).copy()
    out = []
    for s in idx:
        vals = base.copy()
        vals[s] += step
        rp = oracle.run(vals, g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz, f64=True)
        vals[s] -= 2 * step
        rm = oracle.run(vals, g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz, f64=True)
        # fixed topology: same active cells (quad orientation near a tie may flip; L uses V only)
        assert rp.n_vertices == rm.n_vertices == w.shape[0]
        np.testing.assert_array_equal(rp.active, rm.active)
        out.append((np.sum(w * rp.V64) - np.sum(w * rm.V64)) / (2 * step))
    return np.array(out)


@pytest.mark.parametrize("make,seed", [(lambda: synth.csg8(20, seed=7), 0),
                                       (lambda: synth.sphere_unit(16), 1)])
def test_oracle_adjoint_matches_central_differences(make, seed):
    g = make()
    r = oracle.run(np.asarray(g.values, np.float64), g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox,
                   g.oy, g.oz, f64=True)
    rng = np.random.default_rng(seed)
    w = rng.normal(size=(r.n_vertices, 3))
    grad = r.backward(w)
    vals = np.asarray(g.values, np.float64)
    safe = np.abs(vals) > 1e-5  # a +-1e-6 step must not change any sign
    nz = np.flatnonzero((grad != 0) & safe)
    zero = np.flatnonzero((grad == 0) & safe)
    idx = np.concatenate([rng.choice(nz, size=min(150, nz.size), replace=False),
                          rng.choice(zero, size=min(20, zero.size), replace=False)])
    fd = _loss_grad_fd(g, w, idx)
    scale = np.abs(grad).max()
    err = np.abs(fd - grad[idx])
    assert np.all(err <= 1e-6 * scale + 1e-4 * np.abs(grad[idx])), err.max() / scale
