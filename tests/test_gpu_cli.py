"""dc_cli --backend hip (the C++ front end over HipDualContouringBackend): OBJ output and the
--benchmark stats block against the oracle on the reference's generators (src/cpp/dc_cli.cpp:49-118:
sphere, box, plane), --save-sdf output against the generators, .sdf input (:31-47) written by the
Python writer, a short .sdf refused, and --close-holes / --remove-floaters (:334-338) against the
oracle's postprocess_mesh."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import PKG
from contouring_hip import io, synth

pytestmark = pytest.mark.gpu

CLI = os.path.join(PKG, "build", "dc_cli")


def read_obj(path):
    V, F = [], []
    with open(path) as f:
        for line in f:
            if line.startswith("v "):
                V.append([float(t) for t in line.split()[1:4]])
            elif line.startswith("f "):
                F.append([int(t) - 1 for t in line.split()[1:5]])
    return np.array(V), np.array(F, dtype=np.int32)


@pytest.mark.parametrize("precision", ["plain-fp64", "fp32-fallback"])
def test_cli_generate_sphere_matches_oracle(tmp_path, precision):
    out = tmp_path / "sphere.obj"
    res = subprocess.run([CLI, "--generate", "sphere", "--grid-size", "64", "-b", "hip",
                          "--precision", precision, "--benchmark", str(out)],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    assert "Backend: hip" in res.stdout
    V, F = read_obj(out)
    ref = oracle.run_grid(synth.cli_generate("sphere", 64))
    assert F.shape[0] == ref.n_faces
    if precision == "plain-fp64":
        np.testing.assert_array_equal(F, ref.F)
    # OBJ text keeps 6 significant digits
    np.testing.assert_allclose(V, ref.V64, rtol=0, atol=2e-6)
    assert f"Vertex Count: {ref.n_vertices}" in res.stdout


def test_cli_rejects_other_backends(tmp_path):
    res = subprocess.run([CLI, "--generate", "sphere", "-g", "16", "-b", "cuda",
                          str(tmp_path / "x.obj")], capture_output=True, text=True, timeout=60)
    assert res.returncode == 1
    assert "-b hip" in res.stderr


def _cli(*args, timeout=120):
    res = subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=timeout)
    return res


@pytest.mark.parametrize("kind", ["box", "plane", "sphere"])
def test_cli_generators_save_sdf_and_mesh(tmp_path, kind):
    sdf, obj = tmp_path / "g.sdf", tmp_path / "m.obj"
    res = _cli("--generate", kind, "-g", 40, "--save-sdf", sdf, "-o", obj, "--benchmark")
    assert res.returncode == 0, res.stderr
    g = io.load_sdf(sdf)
    ref_grid = synth.cli_generate(kind, 40)
    np.testing.assert_array_equal(g.values, np.asarray(ref_grid.values, np.float32))
    assert (g.nx, g.ny, g.nz, np.float32(g.vx), g.ox) == (40, 40, 40, np.float32(ref_grid.vx), -1.0)
    ref = oracle.run_grid(ref_grid)
    V, F = read_obj(obj)
    np.testing.assert_array_equal(F, ref.F)
    np.testing.assert_allclose(V, ref.V64, rtol=0, atol=2e-6)
    assert f"Face Count: {ref.n_faces}" in res.stdout


def test_cli_reads_sdf_from_the_python_writer(tmp_path):
    g = synth.csg8(64, seed=3)
    sdf, obj = tmp_path / "csg.sdf", tmp_path / "m.obj"
    io.save_sdf(sdf, g)
    res = _cli("-i", sdf, "-o", obj)
    assert res.returncode == 0, res.stderr
    ref = oracle.run_grid(g)
    V, F = read_obj(obj)
    np.testing.assert_array_equal(F, ref.F)
    np.testing.assert_allclose(V, ref.V64, rtol=0, atol=2e-6)
    short = tmp_path / "short.sdf"
    short.write_bytes(sdf.read_bytes()[:-4])
    res = _cli("-i", short, "-o", tmp_path / "x.obj")
    assert res.returncode == 1 and "Failed to read SDF" in res.stderr


@pytest.mark.parametrize("flags", [("--close-holes",), ("--remove-floaters",),
                                   ("--close-holes", "--remove-floaters")])
def test_cli_postprocess_flags(tmp_path, flags):
    # the plane z = const meets the grid's sides: an open sheet with one boundary loop
    obj = tmp_path / "m.obj"
    res = _cli("--generate", "plane", "-g", 40, *flags, "--benchmark", "-o", obj)
    assert res.returncode == 0, res.stderr
    ref = oracle.run_grid(synth.cli_generate("plane", 40))
    V2, F2, info = oracle.postprocess(ref.V32, ref.F, "--close-holes" in flags,
                                      "--remove-floaters" in flags)
    V, F = read_obj(obj)
    np.testing.assert_array_equal(F, F2)
    np.testing.assert_allclose(V, V2, rtol=0, atol=2e-6)
    assert f"Vertex Count: {V2.shape[0]}" in res.stdout
    assert f"Face Count: {F2.shape[0]}" in res.stdout
    if "--close-holes" in flags:
        assert info["loops_filled"] == 1


def _obj_grid(V, res):
    """The grid dc_cli samples a mesh on (the reference's dc_cli.cpp:157-182), in FP32."""
    f32 = np.float32
    lo, hi = V.min(0).astype(f32), V.max(0).astype(f32)
    d = hi - lo
    diag = f32(np.sqrt(f32(f32(f32(d[0] * d[0]) + f32(d[1] * d[1])) + f32(d[2] * d[2]))))
    pad = f32(diag * f32(0.08)) if diag > 0 else f32(0.1)
    lo, hi = (lo - pad).astype(f32), (hi + pad).astype(f32)
    spacing = ((hi - lo) / f32(res - 1)).astype(f32)
    return (res, res, res), tuple(float(x) for x in lo), tuple(float(x) for x in spacing)


@pytest.mark.parametrize("voxelize", [False, True])
def test_cli_mesh_input(tmp_path, voxelize):
    """-i mesh.obj: quads written as faces (split into a fan from their first vertex), tokens with
    texture/normal indices and negative (relative) indices, sampled on the padded bounding-box
    grid by the GPU mesh-to-SDF pass and extracted; equal to the oracle's mesh-to-SDF followed by
    the oracle's extraction. --save-sdf writes the sampled grid."""
    V, T = synth.torus_mesh(R=0.3, r=0.12, nu=40, nv=20)
    nv = len(V)
    lines = [f"v {x:.9g} {y:.9g} {z:.9g}" for x, y, z in V]
    Q = T.reshape(-1, 2, 3)  # (a, b, c), (a, c, d) -> quad (a, b, c, d)
    for q, (t0, t1) in enumerate(Q):
        a, b, c, d = t0[0], t0[1], t0[2], t1[2]
        if q % 3 == 0:
            lines.append(f"f {a + 1}/1/1 {b + 1}//2 {c + 1}/3 {d + 1}")
        elif q % 3 == 1:
            lines.append(f"f {a - nv} {b - nv} {c - nv} {d - nv}")
        else:
            lines.append(f"f {a + 1} {b + 1} {c + 1} {d + 1}")
    obj_in = tmp_path / "torus.obj"
    obj_in.write_text("\n".join(lines) + "\n")
    res_n = 48
    dims, origin, spacing = _obj_grid(V, res_n)
    out, sdf = tmp_path / "out.obj", tmp_path / "grid.sdf"
    args = ["-i", obj_in, "-o", out, "-g", res_n, "--benchmark"]
    if voxelize:
        args += ["--voxelize-first", "--voxel-res", 40]
    r = _cli(*args)
    assert r.returncode == 0, r.stderr
    want = oracle.mesh_to_sdf(V, T, dims, origin, spacing, voxelize, 40)
    ref = oracle.run(want.ravel(), *dims, *spacing, *origin)
    Vo, Fo = read_obj(out)
    np.testing.assert_array_equal(Fo, ref.F)
    np.testing.assert_allclose(Vo, ref.V64, rtol=0, atol=2e-6)
    # the same through the host path with the grid written out
    r = _cli(*args, "--save-sdf", sdf)
    assert r.returncode == 0, r.stderr
    g = io.load_sdf(sdf)
    assert (g.nx, g.ny, g.nz) == dims
    assert np.array_equal(np.asarray(g.values, np.float32).view(np.uint32),
                          want.ravel().view(np.uint32))
    Vo2, Fo2 = read_obj(out)
    np.testing.assert_array_equal(Fo2, ref.F)


def test_cli_mesh_input_errors(tmp_path):
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 7\n")
    r = _cli("-i", bad, "-o", tmp_path / "x.obj", "-g", 16)
    assert r.returncode == 1 and "outside" in r.stderr
    empty = tmp_path / "empty.obj"
    empty.write_text("v 0 0 0\n")
    r = _cli("-i", empty, "-o", tmp_path / "x.obj", "-g", 16)
    assert r.returncode == 1 and "Failed to convert OBJ" in r.stderr


@pytest.mark.parametrize("backend,mode,brick", [("hip-sparse", "finite-difference", 8),
                                                ("hip-sparse-mvdc", "edge-gradient", 4),
                                                ("mvdc", "precomputed-gradient", 3)])
def test_cli_sparse_backends(tmp_path, backend, mode, brick):
    """-b hip-sparse / hip-sparse-mvdc / mvdc with --brick-size and --normal-mode
    (src/cpp/dc_cli.cpp:311-321): OBJ and the stats block against the oracle's sparse backends."""
    out = tmp_path / "m.obj"
    r = _cli("--generate", "box", "-g", 40, "-b", backend, "--normal-mode", mode, "--brick-size",
             brick, "--chunk-size", 1000, "--benchmark", "-o", out)
    assert r.returncode == 0, r.stderr
    mv = backend != "hip-sparse"
    modes = {"finite-difference": 0, "edge-gradient": 1, "precomputed-gradient": 2}
    want = oracle.sparse_grid(synth.cli_generate("box", 40), brick_size=brick,
                              normal_mode=modes[mode], multi_vertex=mv)
    V, F = read_obj(out)
    np.testing.assert_array_equal(F, want["faces"])
    np.testing.assert_allclose(V, want["vertices"], rtol=0, atol=2e-6)
    assert ("Backend: hip-sparse-mvdc" if mv else "Backend: hip-sparse") in r.stdout
    st = want["stats"]
    assert f"Active Bricks: {st['active_bricks']} " in r.stdout
    assert f"Vertex Count: {st['vertex_count']}" in r.stdout
    assert f"Face Count: {st['face_count']}" in r.stdout
    if st["two_vertex_cells"]:
        assert f"Two-Vertex Cells: {st['two_vertex_cells']}" in r.stdout


def test_cli_mesh_input_sparse_mvdc_and_postprocess(tmp_path):
    """The web app's call (webapp/server.py:63-78): dc_cli -i in.obj -o out.obj -g res -b <mvdc>
    --brick-size 8 with the post-processing flags: mesh-to-SDF on the GPU, the multi-vertex
    extraction on the device grid, then the GPU post-processing of the sparse mesh; equal to the
    oracle's three steps."""
    V, T = synth.torus_mesh(R=0.3, r=0.12, nu=40, nv=20)
    obj_in = tmp_path / "torus.obj"
    obj_in.write_text("\n".join([f"v {x:.9g} {y:.9g} {z:.9g}" for x, y, z in V] +
                                [f"f {a + 1} {b + 1} {c + 1}" for a, b, c in T]) + "\n")
    res_n = 48
    dims, origin, spacing = _obj_grid(V, res_n)
    grid = oracle.mesh_to_sdf(V, T, dims, origin, spacing)
    want = oracle.sparse(grid.ravel(), *dims, *spacing, *origin, brick_size=8, multi_vertex=True)
    out = tmp_path / "out.obj"
    r = _cli("-i", obj_in, "-o", out, "-g", res_n, "-b", "hip-sparse-mvdc", "--brick-size", 8)
    assert r.returncode == 0, r.stderr
    Vo, Fo = read_obj(out)
    np.testing.assert_array_equal(Fo, want["faces"])
    np.testing.assert_allclose(Vo, want["vertices"], rtol=0, atol=2e-6)
    V2, F2, _ = oracle.postprocess(want["vertices"], want["faces"], True, True)
    r = _cli("-i", obj_in, "-o", out, "-g", res_n, "-b", "hip-sparse-mvdc", "--close-holes",
             "--remove-floaters", "--benchmark")
    assert r.returncode == 0, r.stderr
    Vo, Fo = read_obj(out)
    np.testing.assert_array_equal(Fo, F2)
    np.testing.assert_allclose(Vo, V2, rtol=0, atol=2e-6)
    assert f"Face Count: {F2.shape[0]}" in r.stdout
