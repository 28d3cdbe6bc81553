"""The NumPy mesh counter (tests/counts.py) against the oracle on the small suite, so that the
full-size GPU tests that rely on it check against a pinned count."""
import pytest

import oracle
from counts import mesh_counts
from grids import small_suite


@pytest.mark.parametrize("idx", range(len(small_suite())))
def test_counts_match_oracle(idx):
    g = small_suite()[idx]
    ref = oracle.run_grid(g)
    # a small slab exercises the slab seams
    assert mesh_counts(g.values, g.nx, g.ny, g.nz, slab=5) == (ref.n_vertices, ref.n_faces), g.name
    assert mesh_counts(g.values, g.nx, g.ny, g.nz) == (ref.n_vertices, ref.n_faces), g.name
