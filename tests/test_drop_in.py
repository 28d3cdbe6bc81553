"""The drop-in boundary compiles inside the reference tree (CPU only).

- dc_hip_backend.h / dc_hip_backend.cpp build against the reference's include/dc_backend.h (and
  include/dc_cuda.h with DC_ENABLE_CUDA) without redefining any of its types;
- the backend object defines none of the symbols the reference's dc_core defines
  (DenseSdfGrid::sample, postprocess_mesh), so it links next to dc_core;
- the reference dc_cli factory with the "hip" branch compiles against the reference headers;
- a pybind11 module built like the reference's _contouring_cpp_module, with bind_hip_backend(m)
  from include/dc_hip_pybind.h, builds, imports and raises RuntimeError on a machine without a
  HIP device.

Reference headers are read from /root/reference/include when that tree is present; the tests that
need it are skipped elsewhere. No reference file is copied. dc_core itself (which defines
DenseSdfGrid::sample in the reference build) needs Eigen and libigl and cannot be built here, so
the module links this repository's standalone definition (csrc/dc_standalone.cpp) in its place.
"""
import os
import subprocess
import sys
import sysconfig

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
CSRC = os.path.join(ROOT, "phanty-sdf-cuda_amd", "csrc")
BUILD = os.path.join(ROOT, "phanty-sdf-cuda_amd", "build")
HERE = os.path.join(ROOT, "tests", "drop_in")
REF_INC = "/root/reference/include"

needs_ref = pytest.mark.skipif(not os.path.exists(os.path.join(REF_INC, "dc_backend.h")),
                               reason="reference headers not present")


def _gxx(args, cwd):
    r = subprocess.run(["g++", "-std=c++17", *args], cwd=cwd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


def test_backend_object_defines_no_reference_symbols(tmp_path):
    obj = str(tmp_path / "dc_hip_backend.o")
    _gxx(["-c", "-fPIC", "-O1", "-I", os.path.join(INC, "standalone"), "-I", INC,
          os.path.join(CSRC, "dc_hip_backend.cpp"), "-o", obj], str(tmp_path))
    syms = subprocess.run(["nm", "-C", "--defined-only", obj], capture_output=True,
                          text=True).stdout
    assert "HipDualContouringBackend::extract" in syms
    for name in ("DenseSdfGrid::sample", "postprocess_mesh"):
        assert name not in syms, name


@needs_ref
@pytest.mark.parametrize("cuda", [False, True])
def test_backend_compiles_against_reference_headers(tmp_path, cuda):
    defs = ["-DDC_ENABLE_CUDA"] if cuda else []
    # the reference's include directory first, as in its CMake target; either order resolves
    # <dc_backend.h> to the reference's, since this repository's include/ no longer holds one
    for order in ([REF_INC, INC], [INC, REF_INC]):
        incs = [x for d in order for x in ("-I", d)]
        _gxx(["-c", "-fPIC", "-O1", *defs, *incs, os.path.join(CSRC, "dc_hip_backend.cpp"),
              "-o", str(tmp_path / "o.o")], str(tmp_path))
        _gxx(["-fsyntax-only", *defs, *incs, os.path.join(HERE, "factory.cpp")], str(tmp_path))


@needs_ref
def test_pybind_module_builds_and_imports(tmp_path):
    pybind = subprocess.run([sys.executable, "-m", "pybind11", "--includes"], capture_output=True,
                            text=True, check=True).stdout.split()
    ext = sysconfig.get_config_var("EXT_SUFFIX")
    mod = str(tmp_path / ("_dc_hip_dropin" + ext))
    if not os.path.exists(os.path.join(BUILD, "libdc_hip.so")):
        pytest.skip("libdc_hip.so not built")
    _gxx(["-shared", "-fPIC", "-O1", *pybind, "-I", REF_INC, "-I", INC,
          os.path.join(HERE, "module.cpp"), os.path.join(CSRC, "dc_hip_backend.cpp"),
          os.path.join(CSRC, "dc_standalone.cpp"), "-L", BUILD, "-ldc_hip",
          "-Wl,-rpath," + BUILD, "-o", mod], str(tmp_path))
    code = (
        "import sys; sys.path.insert(0, sys.argv[1])\n"
        "import _dc_hip_dropin as m\n"
        "g = m.DenseSdfGrid(); g.nx = g.ny = g.nz = 2; g.values = [1.0] * 8\n"
        "d = m.DenseSdfGridDevice(); d.d_values = 4096; assert d.d_values == 4096\n"
        "try:\n"
        "    be = m.HipDualContouringBackend(0, 0)\n"
        "    mesh, st = be.extract(g)\n"
        "    print('extracted', st.vertex_count, len(mesh.faces))\n"
        "except RuntimeError as e:\n"
        "    print('raised', e)\n"
        "try:\n"
        "    sp = m.HipSparseMvdcDualContouringBackend(brick_size=4, normal_mode=1)\n"
        "    assert sp.multi_vertex_cells and sp.brick_size == 4\n"
        "    mesh, st = sp.extract(g)\n"
        "    print('sparse', st.backend, st.vertex_count)\n"
        "except RuntimeError as e:\n"
        "    print('sparse raised', e)\n"
    )
    r = subprocess.run([sys.executable, "-c", code, str(tmp_path)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out, sp = r.stdout.strip().split("\n")
    # without a device the constructor raises; with one, the all-positive grid has no vertices
    assert out.startswith("raised HipDualContouringBackend: no HIP device") or \
        out == "extracted 0 0", out
    assert sp.startswith("sparse raised HipSparseDualContouringBackend: no HIP device") or \
        sp == "sparse hip-sparse-mvdc 0", sp
