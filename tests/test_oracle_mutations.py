"""Each plausible mistake in the oracle is caught by a test of tests/test_oracle_pinned.py (CPU).

A copy of oracle/dc_oracle.cpp with one mistake planted is built into a temporary library, and the
test that names that mistake is run against it in a subprocess (ORACLE_LIB): it must fail. The
unmodified copy, built the same way, must pass all of them (the control). Every mutation here
fails an assertion (none crashes the mutated library).
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "dc_oracle.cpp")
PINNED = os.path.join(ROOT, "tests", "test_oracle_pinned.py")

# (name, text in the oracle, replacement, test that must fail)
MUTATIONS = [
    ("trilinear_weight", "(S101 - S001) * omv * w", "(S101 - S001) * omu * w",
     "test_trilinear_field_hermite_points_and_normals_in_closed_form"),
    ("both_endpoints_inside",
     "if (!inside(i0, j0, k0) && !inside(i1, j1, k1)) return 0;",
     "if (!inside(i0, j0, k0) || !inside(i1, j1, k1)) return 0;",
     "test_edge_with_one_border_endpoint_is_averaged"),
    ("average_active_only", "if (vids[q] < 0) return 0;", "if (vids[q] < 0) vids[q] = v0;",
     "test_zero_endpoint_edge_averaged_only_with_four_active_cells"),
    ("quad_order", "std::sort(keys.begin(), keys.end());", "std::sort(keys.rbegin(), keys.rend());",
     "test_quad_order_and_orientation_brute_force"),
    ("gradient_border_step", "const double dx = coord(g.ox, ip, g.vx) - coord(g.ox, im, g.vx);",
     "const double dx = 2.0 * g.vx;", "test_quad_order_and_orientation_brute_force"),
]


def _build(tmp_path, name, old=None, new=None):
    src = open(SRC).read()
    if old is not None:
        assert src.count(old) == 1, (name, src.count(old))
        src = src.replace(old, new)
    d = tmp_path / name
    d.mkdir()
    (d / "dc_oracle.cpp").write_text(src)
    lib = str(d / "liboracle_dc.so")
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math",
                        "-fopenmp", "-fPIC", "-shared", "-o", lib, str(d / "dc_oracle.cpp")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return lib


def _run_pinned(lib, test=None):
    target = PINNED if test is None else f"{PINNED}::{test}"
    env = dict(os.environ, ORACLE_LIB=lib)
    return subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                           target], cwd=ROOT, env=env, capture_output=True, text=True,
                          timeout=300)


def test_unmodified_copy_passes(tmp_path):
    r = _run_pinned(_build(tmp_path, "control"))
    assert r.returncode == 0, r.stdout[-2000:]


@pytest.mark.parametrize("name,old,new,test", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_mutation_is_caught(tmp_path, name, old, new, test):
    r = _run_pinned(_build(tmp_path, name, old, new), test)
    assert r.returncode != 0, f"mutation {name} survived {test}:\n{r.stdout[-1500:]}"
