"""Build recipe for the MI355X dual-contouring backend (run on the CPU; hipcc cross-compiles).

Outputs (in-tree, git-ignored, shipped with the tree):
  build/libdc_hip.so    HIP kernels for gfx950 + the C ABI (include/dc_hip.h)
  build/libdc_synth.so  synthetic SDF grid generators (host, OpenMP)
  build/dc_cli          the command-line front end (--backend hip)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "build")
INCLUDE = os.path.join(ROOT, "include")
# the restated reference types (dc_backend.h) for this repository's own C++ builds; a build inside
# the reference tree uses the reference's include/dc_backend.h instead
STANDALONE = os.path.join(INCLUDE, "standalone")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops: no packed FP32 VALU ops (some of their forms have misbehaved on these boxes).
HIP_FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-ffp-contract=off", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I", INCLUDE, "-I", CSRC, "-Wno-unused-command-line-argument"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ...")
    return res


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in ("dc_kernels.hip", "dc_cells.hip", "dc_cells_fast.hip", "dc_post.hip",
                "dc_m2s.hip", "dc_sparse.hip", "dc_hip.cpp"):
        obj = os.path.join(BUILD, src.rsplit(".", 1)[0] + ".o")
        flags = list(HIP_FLAGS)
        if src == "dc_cells_fast.hip":  # results held to tolerances: FMA contraction allowed
            flags[flags.index("-ffp-contract=off")] = "-ffp-contract=fast"
        _run([HIPCC, *flags, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    lib = os.path.join(BUILD, "libdc_hip.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", lib, *objs], verbose)
    _run(["g++", "-std=c++17", "-O3", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared",
          "-o", os.path.join(BUILD, "libdc_synth.so"), os.path.join(CSRC, "synth.cpp")], verbose)
    cli_srcs = [os.path.join(CSRC, f)
                for f in ("dc_cli.cpp", "dc_hip_backend.cpp", "dc_standalone.cpp", "synth.cpp")]
    _run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fopenmp", "-I", STANDALONE,
          "-I", INCLUDE, "-I", CSRC,
          "-o", os.path.join(BUILD, "dc_cli"), *cli_srcs, "-L", BUILD, "-ldc_hip",
          "-Wl,-rpath,$ORIGIN"], verbose)
    return lib


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
