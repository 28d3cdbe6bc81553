"""The list-entry decoders that bound every per-cell and per-face kernel (csrc/dc_decode.h), on the
CPU: valid entries decode to their cells and quad edges; a corrupted entry (a padding bit past cx,
a word outside the grid's plane, a face bit past cz or on an edge that emits no quad) is rejected,
so the kernels skip it and flag DcCounters::error (reported as DC_HIP_ERR_INTERNAL) instead of
loading samples past the grid."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "phanty-sdf-cuda_amd", "csrc")


def test_decoders_accept_valid_and_reject_corrupted_entries(tmp_path):
    exe = str(tmp_path / "decode_check")
    src = os.path.join(ROOT, "tests", "decode", "decode_check.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", CSRC, src, "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout
