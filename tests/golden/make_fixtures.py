"""Writes tests/golden/oracle_fixtures.npz: small grids and the oracle's outputs for them.

The reference ships no golden vectors; these freeze the outputs of the oracle after it has been
checked against the independent NumPy twin, closed forms and library eigen-solvers
(tests/test_oracle.py), so later changes to the oracle or the kernels are caught.
Run from the repository root: python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "phanty-sdf-cuda_amd"), os.path.dirname(HERE)]

import oracle  # noqa: E402
from grids import ragged_sphere, ref_sphere, ref_thin_gap, zeros_grid  # noqa: E402
from contouring_hip import synth  # noqa: E402


def main():
    grids = [ref_sphere(16), ref_thin_gap(16), zeros_grid(), ragged_sphere(13, 7, 9),
             synth.csg8(16, seed=1234), synth.sphere_unit(16)]
    out = {}
    for g in grids:
        r = oracle.run_grid(g)
        name = g.name
        out[f"{name}__values"] = np.asarray(g.values, np.float32)
        out[f"{name}__dims"] = np.array([g.nx, g.ny, g.nz], np.int32)
        out[f"{name}__geo"] = np.array([g.vx, g.vy, g.vz, g.ox, g.oy, g.oz], np.float32)
        out[f"{name}__faces"] = r.F
        out[f"{name}__active"] = r.active
        out[f"{name}__v64"] = r.V64
        print(name, r.n_vertices, r.n_faces)
    np.savez_compressed(os.path.join(HERE, "oracle_fixtures.npz"), **out)


if __name__ == "__main__":
    main()
