"""Batched extraction (dc_hip_extract_batch) on the GPU.

Each grid's slice of a batch must be bit-identical to the single-grid extraction of that grid
(vertex ids are local to their grid) and, for the plain FP64 mode, match the oracle like a single
extraction does: quads exact, FP32 vertices exact, FP64 vertices within 1e-6*h.
"""
import numpy as np
import pytest

import oracle
from grids import small_suite
from contouring_hip import DenseSdfGridDevice, HipDualContouringBackend, synth

pytestmark = pytest.mark.gpu


def _device(grids):
    import torch
    out = []
    for g in grids:
        phi = torch.from_numpy(np.ascontiguousarray(g.values, np.float32)).cuda()
        out.append(DenseSdfGridDevice(g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz,
                                      d_values=phi))
    return out


def _batch_suite():
    # uneven sizes, empty grids (constant sign) in the middle, a 2^3 grid, non-cubic grids
    return small_suite() + [synth.csg8(72, seed=3), synth.csg8(24, seed=8)]


@pytest.mark.parametrize("precision", ["plain-fp64", "fp32-fallback"])
def test_batch_slices_equal_single_extractions(precision):
    import torch
    grids = _batch_suite()
    dev = _device(grids)
    be = HipDualContouringBackend(precision)
    print(f"[batch] {precision}: {len(grids)} grids", flush=True)
    pairs, st = be.extract_batch(dev, with_f64=(precision == "plain-fp64"))
    assert len(pairs) == len(grids)
    assert st.vertex_count == sum(int(p[0].shape[0]) for p in pairs)
    assert st.face_count == sum(int(p[1].shape[0]) for p in pairs)
    T_all = be.triangles()  # grid-local ids, 2 triangles per quad, grid-major
    single = HipDualContouringBackend(precision)
    f0 = 0
    for g, d, item in zip(grids, dev, pairs):
        (V, F), _ = single.extract_tensor(d.d_values.reshape(g.nz, g.ny, g.nx),
                                          (g.vx, g.vy, g.vz), (g.ox, g.oy, g.oz))
        assert torch.equal(item[0], V), g.name
        assert torch.equal(item[1], F), g.name
        nf = int(F.shape[0])
        assert torch.equal(T_all[2 * f0:2 * (f0 + nf)], single.triangles()), g.name
        f0 += nf


def test_batch_plain_matches_oracle():
    grids = _batch_suite()
    dev = _device(grids)
    be = HipDualContouringBackend("plain-fp64")
    pairs, _ = be.extract_batch(dev, with_f64=True)
    for g, (V, F, V64) in zip(grids, pairs):
        print(f"[batch] oracle {g.name}", flush=True)
        ref = oracle.run_grid(g)
        assert V.shape[0] == ref.n_vertices, g.name
        np.testing.assert_array_equal(F.cpu().numpy(), ref.F.reshape(-1, 4), err_msg=g.name)
        if ref.n_vertices:
            np.testing.assert_array_equal(V.cpu().numpy(), ref.V32, err_msg=g.name)
            h = min(g.vx, g.vy, g.vz)
            assert np.abs(V64.cpu().numpy() - ref.V64).max() <= 1e-6 * h, g.name


def test_batch_repeated_grid_and_reuse():
    """The same grid twice in one batch gives equal slices; a second, smaller batch on the same
    context reuses its buffers and still matches."""
    import torch
    g1, g2 = synth.csg8(40, seed=2), synth.sphere_unit(33)
    dev = _device([g1, g2, g1])
    be = HipDualContouringBackend("plain-fp64")
    pairs, _ = be.extract_batch(dev)
    assert torch.equal(pairs[0][0], pairs[2][0]) and torch.equal(pairs[0][1], pairs[2][1])
    pairs2, _ = be.extract_batch(dev[1:2])
    assert torch.equal(pairs2[0][0], pairs[1][0]) and torch.equal(pairs2[0][1], pairs[1][1])
