"""The C ABI without a GPU: the library loads, exports every symbol include/dc_hip.h declares,
and reports a missing device through its status codes (never exits)."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

from conftest import ROOT, gpu_available
from contouring_hip import _lib

HEADER = os.path.join(ROOT, "include", "dc_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(dc_hip_[a-z_]+)\s*\(", text)))


def test_header_declares_what_binding_expects():
    assert declared_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = _lib.load_hip()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.lib_path()], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (dc_hip_\w+)", out))
    assert set(declared_symbols()) <= exported


def test_header_is_plain_c():
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-x", "c", HEADER],
                   check=True)


def test_status_strings():
    lib = _lib.load_hip()
    for code in range(9):
        assert lib.dc_hip_status_string(code)
    assert lib.dc_hip_status_string(99) == b"unknown status"


@pytest.mark.skipif(gpu_available(), reason="checks the no-device path")
def test_create_without_device_reports_no_device():
    lib = _lib.load_hip()
    ctx = ctypes.c_void_p()
    rc = lib.dc_hip_create(0, ctypes.byref(ctx))
    assert rc == _lib.DC_HIP_ERR_NO_DEVICE
    assert not ctx.value


def test_null_arguments_are_rejected():
    lib = _lib.load_hip()
    assert lib.dc_hip_create(0, None) == _lib.DC_HIP_ERR_INVALID_ARG
    assert lib.dc_hip_set_option(None, b"timing", 1) == _lib.DC_HIP_ERR_INVALID_ARG
    assert lib.dc_hip_extract(None, None, 0, None) == _lib.DC_HIP_ERR_INVALID_ARG
    assert lib.dc_hip_result_counts(None, None, None) == _lib.DC_HIP_ERR_INVALID_ARG
    assert lib.dc_hip_last_error(None) == b"no context"


def test_backend_raises_without_device():
    if gpu_available():
        pytest.skip("checks the no-device path")
    from contouring_hip import HipDualContouringBackend
    with pytest.raises(_lib.DcHipError):
        HipDualContouringBackend()


def _c_layout(struct, fields, tmp_path):
    """sizeof and offsetof of a dc_hip.h struct, from a C program built against the header."""
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "dc_hip.h"', "int main(void) {",
             f'  printf("%zu\\n", sizeof({struct}));']
    lines += [f'  printf("%zu\\n", offsetof({struct}, {f}));' for f in fields]
    lines += ["  return 0;", "}"]
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), "-o", str(exe), str(src)],
                   check=True)
    return [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]


@pytest.mark.parametrize("struct,mirror", [("dc_hip_stats", _lib.DcHipStats),
                                           ("dc_hip_grid", _lib.DcHipGrid)])
def test_ctypes_mirrors_match_the_header(struct, mirror, tmp_path):
    names = [f[0] for f in mirror._fields_]
    got = _c_layout(struct, names, tmp_path)
    assert got[0] == ctypes.sizeof(mirror)
    for name, off in zip(names, got[1:]):
        assert getattr(mirror, name).offset == off, name


def test_autotune_rejects_bad_arguments():
    lib = _lib.load_hip()
    assert lib.dc_hip_autotune(None, None, 0, 0, 1, None, None, None) == _lib.DC_HIP_ERR_INVALID_ARG


def test_one_hip_runtime_per_process():
    """Loaded from Python before torch has touched the device, the library shares torch's HIP
    runtime: a second runtime beside it would leave whichever initialized the GPU second with no
    device."""
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from contouring_hip import _lib\n"
            "_lib.load_hip()\n"
            "import torch\n"
            "maps = open('/proc/self/maps').read().splitlines()\n"
            "print(len({l.split()[-1] for l in maps if 'libamdhip64' in l}))\n"
            % os.path.join(ROOT, "phanty-sdf-cuda_amd"))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                         timeout=300, check=True).stdout
    assert out.strip().splitlines()[-1] == "1"
