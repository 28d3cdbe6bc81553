"""The two marking passes at their block boundaries, against the oracle.

The sign pass packs 64 sample layers per entry and 64 samples per x word; the cell pass reads the
next block's bit 0 for a cell's upper layer and the next word's pencil for its x + 1 corner.
These grids put the surface across those boundaries: nz = 65 (the next block holds one layer),
nx = 65 and 129 (the last cell's x + 1 corner is the first pencil of the next sign word), and
ny values that are not multiples of the row tile. Every sign-pass tile shape must give the same
mesh, and so must the whole-row form of the sign pass (8 samples per lane, rows of 8k samples).
"""
import numpy as np
import pytest

import oracle
from grids import ragged_sphere
from contouring_hip import HipDualContouringBackend

pytestmark = pytest.mark.gpu

# (nx, ny, nz, sphere radius): thin grids need a larger sphere to cross them
GRIDS = [(65, 17, 65, 0.4), (129, 33, 66, 0.4), (65, 70, 129, 0.4), (3, 2, 130, 0.6),
         (130, 3, 2, 0.6), (5, 4, 193, 0.5)]


def _extract(be, g):
    import torch
    phi = torch.from_numpy(np.ascontiguousarray(g.values, np.float32)).cuda().reshape(
        g.nz, g.ny, g.nx)
    return be.extract_tensor(phi, (g.vx, g.vy, g.vz), (g.ox, g.oy, g.oz), with_f64=True)


@pytest.mark.parametrize("dims", GRIDS)
def test_block_boundaries_match_oracle(dims):
    g = ragged_sphere(*dims[:3], r=dims[3])
    print(f"[marking] {g.name}", flush=True)
    be = HipDualContouringBackend("plain-fp64")
    (V, F, V64), st = _extract(be, g)
    ref = oracle.run_grid(g)
    assert st.vertex_count == ref.n_vertices and st.face_count == ref.n_faces
    np.testing.assert_array_equal(F.cpu().numpy(), ref.F.reshape(-1, 4))
    if ref.n_vertices:
        np.testing.assert_array_equal(V.cpu().numpy(), ref.V32)
        h = min(g.vx, g.vy, g.vz)
        assert np.abs(V64.cpu().numpy() - ref.V64).max() <= 1e-6 * h


@pytest.mark.parametrize("dims", [(65, 17, 65), (129, 33, 66)])
def test_every_tile_shape_gives_the_same_mesh(dims):
    import torch
    g = ragged_sphere(*dims, r=0.4)
    ref = None
    for rows in (4, 8, 16, 32):
        for waves in (1, 2, 4):
            print(f"[marking] {g.name} rows {rows} waves {waves}", flush=True)
            be = HipDualContouringBackend("plain-fp64", mark_rows=rows, mark_waves=waves)
            (V, F, V64), _ = _extract(be, g)
            if ref is None:
                ref = (V, F, V64)
            else:
                assert torch.equal(V, ref[0]) and torch.equal(F, ref[1])
                assert torch.equal(V64, ref[2])


# rows of 8k samples: the whole-row sign pass applies (one or two 512-sample row segments)
WIDE_GRIDS = [(72, 17, 65, 0.4), (520, 9, 66, 0.5), (64, 64, 130, 0.4), (8, 3, 70, 0.6)]


@pytest.mark.parametrize("dims", WIDE_GRIDS)
def test_whole_row_sign_pass_matches_oracle_and_tiles(dims):
    import torch
    g = ragged_sphere(*dims[:3], r=dims[3])
    print(f"[marking] whole rows {g.name}", flush=True)
    ref = oracle.run_grid(g)
    got = {}
    for rows in (1, 16, 0):
        be = HipDualContouringBackend("plain-fp64", mark_rows=rows)
        (V, F, V64), st = _extract(be, g)
        assert st.vertex_count == ref.n_vertices and st.face_count == ref.n_faces, rows
        got[rows] = (V, F, V64)
    np.testing.assert_array_equal(got[1][1].cpu().numpy(), ref.F.reshape(-1, 4))
    if ref.n_vertices:
        np.testing.assert_array_equal(got[1][0].cpu().numpy(), ref.V32)
    for rows in (16, 0):
        assert all(torch.equal(a, b) for a, b in zip(got[rows], got[1])), rows


def test_whole_row_sign_pass_falls_back_for_unaligned_values():
    """Values that start 4 bytes past a 16-byte boundary cannot take the whole-row form's 16-byte
    loads: the 16-row form runs instead, with the oracle's mesh."""
    import torch
    g = ragged_sphere(72, 17, 65, r=0.4)
    n = g.nx * g.ny * g.nz
    base = torch.zeros(n + 4, dtype=torch.float32, device="cuda")
    base[1:1 + n] = torch.from_numpy(np.ascontiguousarray(g.values, np.float32)).cuda()
    view = base[1:1 + n]
    assert view.data_ptr() % 16 == 4
    be = HipDualContouringBackend("plain-fp64", mark_rows=1)
    (V, F, V64), st = be.extract_tensor(view.reshape(g.nz, g.ny, g.nx), (g.vx, g.vy, g.vz),
                                        (g.ox, g.oy, g.oz), with_f64=True)
    ref = oracle.run_grid(g)
    assert st.vertex_count == ref.n_vertices and st.face_count == ref.n_faces
    np.testing.assert_array_equal(F.cpu().numpy(), ref.F.reshape(-1, 4))
    np.testing.assert_array_equal(V.cpu().numpy(), ref.V32)
