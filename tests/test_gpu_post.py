"""Mesh post-processing on the GPU (dc_hip_postprocess) against the oracle's restatement of the
reference's postprocess_mesh (src/cpp/dc_backend.cpp:217-558): vertices and faces equal array for
array, every flag combination, on hand-built meshes (a quad, two quads sharing a vertex, a book of
three quads on one edge, regular-polygon frustums, one with loops too long to fill), random
welded soups, DC meshes of disjoint spheres and of a sphere cut by the grid, csg8 at 128^3 and
512^3 (the last extraction's mesh, straight from the device), device tensors in and out, the error
paths, and the counts reported beside the oracle's."""
import numpy as np
import pytest

import oracle
from contouring_hip import DualContouringMesh, HipDualContouringBackend, synth
from contouring_hip._lib import DcHipError, DC_HIP_ERR_INVALID_ARG, DC_HIP_ERR_NO_FORWARD

pytestmark = pytest.mark.gpu

FLAGS = [(True, True), (True, False), (False, True), (False, False)]


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def be():
    return HipDualContouringBackend("plain-fp64")


def _check(be, V, F, close_holes, remove_floaters, label):
    _say(f"[post] {label} close={close_holes} floaters={remove_floaters} "
         f"({V.shape[0]} vertices, {F.shape[0]} quads)")
    out, st = be.postprocess(DualContouringMesh(V.ravel(), F.ravel()), close_holes, remove_floaters)
    V2, F2, info = oracle.postprocess(V, F, close_holes, remove_floaters)
    np.testing.assert_array_equal(out.faces.reshape(-1, 4), F2, err_msg=label)
    np.testing.assert_array_equal(out.vertices.reshape(-1, 3), V2, err_msg=label)
    if close_holes or remove_floaters:
        for k, v in info.items():
            assert st[k] == v, (label, k, st[k], v)
    assert (st["vertex_count"], st["face_count"]) == (V2.shape[0], F2.shape[0])
    return out, st


def _frustum(n, r0=1.0, r1=2.0):
    ang = 2 * np.pi * np.arange(n) / n
    inner = np.stack([r0 * np.cos(ang), r0 * np.sin(ang), np.zeros(n)], 1)
    outer = np.stack([r1 * np.cos(ang), r1 * np.sin(ang), np.ones(n)], 1)
    V = np.concatenate([inner, outer]).astype(np.float32)
    F = np.array([[k, (k + 1) % n, n + (k + 1) % n, n + k] for k in range(n)], np.int32)
    return V, F


def _small_meshes():
    q = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    yield "quad", q, np.array([[0, 1, 2, 3]], np.int32)
    V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [2, 1, 0], [2, 2, 0], [1, 2, 0]],
                 np.float32)
    yield "bowtie", V, np.array([[0, 1, 2, 3], [2, 4, 5, 6]], np.int32)
    V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [1, -1, 0], [0, -1, 0], [1, 0, 1],
                  [0, 0, 1]], np.float32)
    yield "book", V, np.array([[0, 1, 2, 3], [1, 0, 5, 4], [0, 1, 6, 7]], np.int32)
    Vb = np.concatenate([q + np.float32(5), q])
    yield "tie", Vb, np.array([[4, 5, 6, 7], [0, 1, 2, 3]], np.int32)
    for n in (3, 4, 5, 8, 37):
        V, F = _frustum(n)
        yield f"frustum{n}", V, F
    V, F = _frustum(5000, 1.0, 1.5)
    yield "frustum5000", V, F


@pytest.mark.parametrize("close_holes,remove_floaters", FLAGS)
def test_small_meshes_equal_oracle(be, close_holes, remove_floaters):
    for label, V, F in _small_meshes():
        _check(be, V, F, close_holes, remove_floaters, label)


@pytest.mark.parametrize("seed", [0, 1])
@pytest.mark.parametrize("close_holes,remove_floaters", FLAGS[:3])
def test_random_soups_equal_oracle(be, seed, close_holes, remove_floaters):
    # quads over few vertices with exact and near duplicates: welds, dropped faces, edges of many
    # faces, tangled boundary walks
    rng = np.random.default_rng(seed)
    base = rng.uniform(-1, 1, (200, 3)).astype(np.float32)
    dup = base[rng.integers(0, 200, 100)]
    V = np.concatenate([base, dup, dup + np.float32(1e-7)])
    F = rng.integers(0, V.shape[0], (600, 4)).astype(np.int32)
    F[:100, 3] = F[:100, 2]
    _check(be, V, F, close_holes, remove_floaters, f"soup{seed}")


def _sphere_field(n, centres_radii):
    ax = np.arange(n, dtype=np.float32) / np.float32(n - 1)
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    d = np.full(x.shape, np.inf, np.float32)
    for c, r in centres_radii:
        d = np.minimum(d, np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) - r)
    return d.astype(np.float32)


def _dc(be, vals, n):
    h = 1.0 / (n - 1)
    R = oracle.run(vals, n, n, n, h, h, h)
    return R.V32, R.F


@pytest.mark.parametrize("close_holes,remove_floaters", FLAGS[:3])
def test_dc_meshes_equal_oracle(be, close_holes, remove_floaters):
    blobs = _sphere_field(48, [((0.25, 0.25, 0.3), 0.12), ((0.65, 0.6, 0.6), 0.2),
                               ((0.2, 0.75, 0.75), 0.08)])
    cut = _sphere_field(40, [((0.5, 0.5, 0.0), 0.45)])
    corner = _sphere_field(33, [((0.0, 0.0, 0.0), 0.6), ((1.0, 1.0, 1.0), 0.3)])
    for label, vals, n in (("blobs48", blobs, 48), ("cut40", cut, 40), ("corners33", corner, 33)):
        V, F = _dc(be, vals, n)
        _check(be, V, F, close_holes, remove_floaters, label)
    g = synth.csg8(128, seed=7)
    R = oracle.run_grid(g)
    _check(be, R.V32, R.F, close_holes, remove_floaters, "csg8_128")


@pytest.mark.timeout(300)
def test_csg8_512_last_result_equals_oracle(be):
    import torch
    from contouring_hip import DenseSdfGridDevice
    g = synth.csg8(512, seed=1234)
    phi = torch.from_numpy(np.asarray(g.values, np.float32)).cuda()
    dg = DenseSdfGridDevice(g.nx, g.ny, g.nz, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz, d_values=phi)
    mesh, st = be.extract_device(dg)
    V = mesh.vertices.reshape(-1, 3)
    F = mesh.faces.reshape(-1, 4)
    _say(f"[post] csg8 512^3: {V.shape[0]} vertices, {F.shape[0]} quads; GPU postprocess")
    out, pst = be.postprocess(None, True, True)  # the last extraction, read on the device
    _say(f"[post] csg8 512^3: GPU {pst['total_ms']:.1f} ms (walk {pst['walk_ms']:.1f} ms); oracle")
    V2, F2, info = oracle.postprocess(V, F, True, True)
    np.testing.assert_array_equal(out.faces.reshape(-1, 4), F2)
    np.testing.assert_array_equal(out.vertices.reshape(-1, 3), V2)
    for k, v in info.items():
        assert pst[k] == v, (k, pst[k], v)
    # the extraction's result is untouched: its counts and an adjoint still work
    assert be.counts() == (V.shape[0], F.shape[0])
    gv = torch.ones(3 * V.shape[0], dtype=torch.float32, device="cuda")
    gphi = torch.empty_like(phi)
    be.backward(gv, gphi)
    torch.cuda.synchronize()
    assert torch.isfinite(gphi).all()


def test_device_tensors_in_and_out(be):
    import torch
    V, F = _frustum(12)
    Vt = torch.from_numpy(V.ravel()).cuda()
    Ft = torch.from_numpy(F.ravel()).cuda()
    out, _ = be.postprocess(DualContouringMesh(Vt, Ft), True, False, device=True)
    assert out.vertices.is_cuda and out.faces.is_cuda
    V2, F2, _ = oracle.postprocess(V, F, True, False)
    np.testing.assert_array_equal(out.vertices.cpu().numpy().reshape(-1, 3), V2)
    np.testing.assert_array_equal(out.faces.cpu().numpy().reshape(-1, 4), F2)


def test_errors(be):
    fresh = HipDualContouringBackend("plain-fp64")
    with pytest.raises(DcHipError) as e:
        fresh.postprocess(None)  # no extraction yet
    assert e.value.status == DC_HIP_ERR_NO_FORWARD
    V = np.zeros((4, 3), np.float32)
    with pytest.raises(DcHipError) as e:
        be.postprocess(DualContouringMesh(V.ravel(), np.array([0, 1, 2, 4], np.int32)))
    assert e.value.status == DC_HIP_ERR_INVALID_ARG
    with pytest.raises(DcHipError) as e:
        be.postprocess(DualContouringMesh(V.ravel(), np.array([0, 1, -1, 3], np.int32)))
    assert e.value.status == DC_HIP_ERR_INVALID_ARG
    # a batch result: face ids are per grid
    import torch
    from contouring_hip import DenseSdfGridDevice
    grids = []
    for n in (16, 20):
        vals = torch.from_numpy(_sphere_field(n, [((0.5, 0.5, 0.5), 0.3)]).ravel()).cuda()
        h = 1.0 / (n - 1)
        grids.append(DenseSdfGridDevice(n, n, n, h, h, h, 0, 0, 0, d_values=vals))
    fresh.extract_batch(grids)
    with pytest.raises(DcHipError) as e:
        fresh.postprocess(None)
    assert e.value.status == DC_HIP_ERR_INVALID_ARG
    # an empty mesh passes through
    out, st = be.postprocess(DualContouringMesh(np.zeros(0, np.float32), np.zeros(0, np.int32)))
    assert out.vertices.size == 0 and out.faces.size == 0
