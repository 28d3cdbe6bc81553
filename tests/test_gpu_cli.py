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
