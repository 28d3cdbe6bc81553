"""dc_cli --backend hip (the C++ front end over HipDualContouringBackend): OBJ output and the
--benchmark stats block, against the oracle on the reference's own generator
(src/cpp/dc_cli.cpp:49-71, `--generate sphere -g 64`)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import PKG
from contouring_hip import synth

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
