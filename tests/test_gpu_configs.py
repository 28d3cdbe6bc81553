"""The benchmark configurations at their full sizes, on the GPU.

- csg8 (8 rotated boxes minus a sphere) at 256^3 and 512^3: full-array parity with the oracle
  (vertex count, every quad, FP32 vertices bit-exact, FP64 vertices within 1e-6*h), and the
  FP32 + FP64-fallback mode's quads equal to the oracle's with vertices within 1e-4*h;
- blend64 (smooth-min of 64 spheres, clamped to +-4h) at 1024^3: the active-cell and quad counts
  against the NumPy count, full parity with the oracle, and the FP32 + FP64-fallback mode's max
  and RMS vertex error against the plain FP64 vertices (tolerance 1e-4*h), its quads equal to the
  plain mode's;
- 256 uneven CSG grids (32^3 ... 256^3) in one batch: every grid's counts against the NumPy
  count and every 16th grid against the oracle, array for array.
"""
import time

import numpy as np
import pytest

import oracle
from counts import mesh_counts
from contouring_hip import DenseSdfGrid, DenseSdfGridDevice, HipDualContouringBackend, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _host_grid(g):
    return DenseSdfGrid(g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz, g.values)


def _full_parity(label, g, ref, V32, V64, F):
    h = min(g.vx, g.vy, g.vz)
    assert V32.shape[0] == ref.n_vertices, label
    assert F.shape[0] == ref.n_faces, label
    np.testing.assert_array_equal(F, ref.F, err_msg=label)
    if ref.n_vertices:
        np.testing.assert_array_equal(V32, ref.V32, err_msg=label)
        err = float(np.abs(V64 - ref.V64).max())
        assert err <= 1e-6 * h, (label, err / h)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("n", [256, 512])
def test_csg8_full_arrays(n):
    _say(f"[configs] csg8 {n}^3: synth")
    g = synth.csg8(n, seed=1234)
    be = HipDualContouringBackend("plain-fp64")
    _say(f"[configs] csg8 {n}^3: extract")
    mesh, st = be.extract(_host_grid(g))
    _say(f"[configs] csg8 {n}^3: oracle ({st.vertex_count} vertices, {st.face_count} quads)")
    t0 = time.perf_counter()
    ref = oracle.run_grid(g)
    _say(f"[configs] csg8 {n}^3: oracle {time.perf_counter() - t0:.1f} s")
    _full_parity(f"csg8_{n}", g, ref, mesh.vertices.reshape(-1, 3),
                 mesh.vertices_f64.reshape(-1, 3), mesh.faces.reshape(-1, 4))
    assert (st.vertex_count, st.face_count) == mesh_counts(g.values, g.nx, g.ny, g.nz)
    # FP32 + FP64-fallback mode: the same quads, vertices within 1e-4 h
    m32, _ = HipDualContouringBackend("fp32-fallback").extract(_host_grid(g))
    np.testing.assert_array_equal(m32.faces.reshape(-1, 4), ref.F)
    err = np.abs(m32.vertices.reshape(-1, 3).astype(np.float64) - ref.V64).max() / g.h
    _say(f"[configs] csg8 {n}^3: fp32 max err {err:.3e} h")
    assert err <= 1e-4


@pytest.mark.timeout(900)
def test_blend64_1024_counts_parity_and_fp32_error():
    import torch
    _say("[configs] blend64 1024^3: synth")
    g = synth.blend64(1024)
    h = g.h
    phi = torch.from_numpy(np.asarray(g.values, np.float32)).cuda()
    dg = DenseSdfGridDevice(g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz, d_values=phi)
    plain = HipDualContouringBackend("plain-fp64")
    _say("[configs] blend64 1024^3: plain extract")
    mesh, st = plain.extract_device(dg)
    V64 = mesh.vertices_f64.reshape(-1, 3)
    F = mesh.faces.reshape(-1, 4)
    _say(f"[configs] blend64 1024^3: {st.vertex_count} vertices, {st.face_count} quads; counts")
    assert (st.vertex_count, st.face_count) == mesh_counts(g.values, g.nx, g.ny, g.nz)
    fast = HipDualContouringBackend("fp32-fallback")
    _say("[configs] blend64 1024^3: fp32-fallback extract")
    m32, st32 = fast.extract_device(dg)
    V32f = m32.vertices.reshape(-1, 3).astype(np.float64)
    assert V32f.shape == V64.shape
    np.testing.assert_array_equal(m32.faces.reshape(-1, 4), F)  # topology exact in FP32 mode
    d = np.abs(V32f - V64).max(axis=1)
    _say(f"[configs] blend64 1024^3: fp32 max err {d.max() / h:.3e} h, rms "
         f"{np.sqrt((d ** 2).mean()) / h:.3e} h, fallback cells {st32.qef_fallback_count}")
    assert d.max() <= 1e-4 * h
    del phi, dg, m32, V32f, d
    _say("[configs] blend64 1024^3: oracle")
    t0 = time.perf_counter()
    ref = oracle.run_grid(g)
    _say(f"[configs] blend64 1024^3: oracle {time.perf_counter() - t0:.1f} s")
    _full_parity("blend64_1024", g, ref, mesh.vertices.reshape(-1, 3), V64, F)


@pytest.mark.timeout(600)
def test_batch_256_uneven_grids():
    import torch
    _say("[configs] batch of 256 uneven grids: synth")
    grids = synth.batch_grids(256)
    dev = []
    for g in grids:
        phi = torch.from_numpy(np.ascontiguousarray(g.values, np.float32)).cuda()
        dev.append(DenseSdfGridDevice(g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz,
                                      d_values=phi))
    be = HipDualContouringBackend("plain-fp64")
    _say(f"[configs] batch: extract ({sum(g.nx * g.ny * g.nz for g in grids)} samples)")
    pairs, st = be.extract_batch(dev, with_f64=True)
    assert len(pairs) == 256
    for q, (g, (V, F, V64)) in enumerate(zip(grids, pairs)):
        counts = mesh_counts(g.values, g.nx, g.ny, g.nz)
        assert (V.shape[0], F.shape[0]) == counts, (q, g.name)
        if q % 16 == 0:
            _say(f"[configs] batch: oracle grid {q} ({g.name})")
            ref = oracle.run_grid(g)
            _full_parity(g.name, g, ref, V.cpu().numpy(), V64.cpu().numpy(), F.cpu().numpy())
    assert st.vertex_count == sum(int(p[0].shape[0]) for p in pairs)
