"""Each plausible mistake in the oracle is caught by a pinning test (CPU).

A copy of the oracle's sources (oracle/dc_oracle.cpp, oracle/post_oracle.cpp,
oracle/m2s_oracle.cpp, oracle/sparse_oracle.cpp) with one mistake planted in one of them is built
into a temporary library, and the test that names that mistake (tests/test_oracle_pinned.py,
tests/test_post_oracle_pinned.py, tests/test_m2s_oracle_pinned.py,
tests/test_sparse_oracle_pinned.py) is run against it in a subprocess
(ORACLE_LIB): it must fail. The unmodified copy, built the same way, must pass all of them (the
control). Every mutation here fails an assertion (none crashes the mutated library).
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SOURCES = ("dc_oracle.cpp", "post_oracle.cpp", "m2s_oracle.cpp", "sparse_oracle.cpp")
PINNED = os.path.join(ROOT, "tests", "test_oracle_pinned.py")
POST_PINNED = os.path.join(ROOT, "tests", "test_post_oracle_pinned.py")
M2S_PINNED = os.path.join(ROOT, "tests", "test_m2s_oracle_pinned.py")
SP_PINNED = os.path.join(ROOT, "tests", "test_sparse_oracle_pinned.py")

# (name, oracle source, text in it, replacement, test file, test that must fail)
MUTATIONS = [
    ("trilinear_weight", "dc_oracle.cpp", "(S101 - S001) * omv * w", "(S101 - S001) * omu * w",
     PINNED, "test_trilinear_field_hermite_points_and_normals_in_closed_form"),
    ("both_endpoints_inside", "dc_oracle.cpp",
     "if (!inside(i0, j0, k0) && !inside(i1, j1, k1)) return 0;",
     "if (!inside(i0, j0, k0) || !inside(i1, j1, k1)) return 0;",
     PINNED, "test_edge_with_one_border_endpoint_is_averaged"),
    ("average_active_only", "dc_oracle.cpp", "if (vids[q] < 0) return 0;",
     "if (vids[q] < 0) vids[q] = v0;",
     PINNED, "test_zero_endpoint_edge_averaged_only_with_four_active_cells"),
    ("quad_order", "dc_oracle.cpp", "std::sort(keys.begin(), keys.end());",
     "std::sort(keys.rbegin(), keys.rend());", PINNED,
     "test_quad_order_and_orientation_brute_force"),
    ("gradient_border_step", "dc_oracle.cpp",
     "const double dx = coord(g.ox, ip, g.vx) - coord(g.ox, im, g.vx);",
     "const double dx = 2.0 * g.vx;", PINNED, "test_quad_order_and_orientation_brute_force"),
    ("weld_eps", "post_oracle.cpp", "diag * 1e-5f", "diag * 1e-4f", POST_PINNED,
     "test_weld_matches_brute_force_keys"),
    ("weld_key_axis", "post_oracle.cpp", "(int64_t)std::round(y * inv)",
     "(int64_t)std::round(x * inv)", POST_PINNED, "test_weld_matches_brute_force_keys"),
    ("weld_keeps_two_distinct", "post_oracle.cpp", "if (distinct < 3) continue;",
     "if (distinct < 2) continue;", POST_PINNED, "test_weld_matches_brute_force_keys"),
    ("component_tie_last", "post_oracle.cpp", "if (best < 0 || size[c] > size[best])",
     "if (best < 0 || size[c] >= size[best])", POST_PINNED,
     "test_edge_of_three_faces_joins_none_and_ties_keep_the_lowest_face"),
    ("nonmanifold_edge_links", "post_oracle.cpp", "} else if (run == 2) {",
     "} else if (run >= 2) {", POST_PINNED,
     "test_edge_of_three_faces_joins_none_and_ties_keep_the_lowest_face"),
    ("loop_not_reversed", "post_oracle.cpp",
     "fill.run(std::vector<int>(walk.rbegin(), walk.rend()));", "fill.run(walk);", POST_PINNED,
     "test_single_quad_becomes_a_closed_pillow"),
    ("ring_step_rule", "post_oracle.cpp", "const float step = d > mean ? mean : d * 0.5f;",
     "const float step = d > mean ? d * 0.5f : mean;", POST_PINNED,
     "test_ring_fill_of_regular_octagons_in_closed_form"),
    ("long_loops_filled", "post_oracle.cpp",
     "if (closed && walk.size() >= 3 && walk.size() <= 4096)",
     "if (closed && walk.size() >= 3)", POST_PINNED, "test_loops_longer_than_4096_stay_open"),
    ("corner_b_region", "m2s_oracle.cpp", "if (d3 >= 0.0f && d4 <= d3) {  // corner b",
     "if (d3 >= 0.0f && d4 >= d3) {  // corner b", M2S_PINNED,
     "test_point_triangle_closest_point"),
    ("normal_flipped", "m2s_oracle.cpp", "if (dot(sub(p, best_q), best_n) < 0.0f) dist = -dist;",
     "if (dot(sub(p, best_q), best_n) > 0.0f) dist = -dist;", M2S_PINNED,
     "test_convex_band_distances_and_signs"),
    ("no_tie_sum", "m2s_oracle.cpp", "} else if (std::fabs(d - best) <= tol) {",
     "} else if (d == best) {", M2S_PINNED, "test_single_bin_signs_match_ray_parity"),
    ("angle_weight", "m2s_oracle.cpp", "if (ang >= 0.0f) w = ang;", "if (ang >= 0.0f) w = len;",
     M2S_PINNED, "test_weighted_normals_in_one_bin"),
    ("bin_size", "m2s_oracle.cpp", "B.gx = std::max(1, in.nx / 8);",
     "B.gx = std::max(1, in.nx / 4);", M2S_PINNED, "test_bins_follow_the_documented_rule"),
    ("margin", "m2s_oracle.cpp", "2.0f * std::fmax(in.vx", "1.0f * std::fmax(in.vx", M2S_PINNED,
     "test_bins_follow_the_documented_rule"),
    ("band", "m2s_oracle.cpp", "const float lim = 2.5f * B.min_dim;",
     "const float lim = 2.0f * B.min_dim;", M2S_PINNED, "test_bins_follow_the_documented_rule"),
    ("cavity_sign", "m2s_oracle.cpp", "if (out[i] == kFar) out[i] = ext[i] ? 10.0f : -10.0f;",
     "if (out[i] == kFar) out[i] = 10.0f;", M2S_PINNED, "test_flood_exterior_and_cavity"),
    ("border_not_seed", "m2s_oracle.cpp",
     "ext[i] = f > 0.0f && (on_border(x, y, z, nx, ny, nz) || f != kFar);",
     "ext[i] = f > 0.0f && f != kFar;", M2S_PINNED, "test_flood_exterior_and_cavity"),
    ("relax_passes", "m2s_oracle.cpp", "for (int pass = 0; pass < 4; ++pass) {",
     "for (int pass = 0; pass < 2; ++pass) {", M2S_PINNED,
     "test_relaxation_column_over_a_flat_face"),
    ("relax_in_place", "m2s_oracle.cpp", "  float* dst = tmp.data();", "  float* dst = out;",
     M2S_PINNED, "test_relaxation_column_over_a_flat_face"),
    ("voxel_reach", "m2s_oracle.cpp", "* 0.8660254f;", "* 0.25f;", M2S_PINNED,
     "test_voxelize_solid"),
    ("no_dilation", "m2s_oracle.cpp", "        grown[x + (int64_t)vres * (y + (int64_t)vres * z)] = any;",
     "        grown[x + (int64_t)vres * (y + (int64_t)vres * z)] =\n"
     "            shell[x + (int64_t)vres * (y + (int64_t)vres * z)];", M2S_PINNED,
     "test_voxelize_growth_seals_a_small_hole"),
    ("jacobi_v_update", "sparse_oracle.cpp", "V[r * 3 + q] = s * vrp + c * vrq;",
     "V[r * 3 + q] = s * vrp - c * vrq;", SP_PINNED,
     "test_jacobi_pinv_matches_eigendecomposition"),
    ("jacobi_transposed", "sparse_oracle.cpp", "w[0] = V[0] * d[0] + V[3] * d[1] + V[6] * d[2];",
     "w[0] = V[0] * d[0] + V[1] * d[1] + V[2] * d[2];", SP_PINNED,
     "test_jacobi_pinv_matches_eigendecomposition"),
    ("jacobi_no_cut", "sparse_oracle.cpp", "if (S[i] > thr) {", "if (S[i] > 0.0f) {", SP_PINNED,
     "test_jacobi_pinv_drops_small_singular_values"),
    ("sparse_trilinear_weight", "sparse_oracle.cpp", "(f[5] - f[4]) * ov * w",
     "(f[5] - f[4]) * ou * w", SP_PINNED, "test_trilinear_normals_away_from_the_cell_centre"),
    ("edge_blend", "sparse_oracle.cpp", "n = add(mul(1.0f - t, ga), mul(t, gb));",
     "n = add(mul(t, ga), mul(1.0f - t, gb));", SP_PINNED,
     "test_edge_gradient_of_a_quadratic_is_the_gradient_at_the_point"),
    ("crossing_t", "sparse_oracle.cpp", "crossing ? fa / (fa - fb + 1e-12f)",
     "crossing ? fb / (fb - fa + 1e-12f)", SP_PINNED,
     "test_crossing_points_are_linear_zeros_of_their_edge"),
    ("brick_either_sign", "sparse_oracle.cpp", "brick[b] = neg && pos;", "brick[b] = neg || pos;",
     SP_PINNED, "test_activation_rule_brute_force"),
    ("brick_extent", "sparse_oracle.cpp", "for (int k = bz * B; k < bz * B + B + 1 && k < g.nz; ++k)",
     "for (int k = bz * B; k < bz * B + B && k < g.nz; ++k)", SP_PINNED,
     "test_activation_rule_brute_force"),
    ("near_rule", "sparse_oracle.cpp", "close = close || fabsf(v) <= near;",
     "close = close || fabsf(v) < 0.5f * near;", SP_PINNED, "test_activation_rule_brute_force"),
    ("thin_threshold", "sparse_oracle.cpp", "return dot(ga, gb) < -0.3f;",
     "return dot(ga, gb) < -0.2f;", SP_PINNED, "test_thin_edges_need_opposed_gradients"),
    ("split_side", "sparse_oracle.cpp", "if (dot(cd.n[q], nA) >= dot(cd.n[q], nB)) {",
     "if (dot(cd.n[q], nA) <= dot(cd.n[q], nB)) {", SP_PINNED, "test_thin_sheet_in_closed_form"),
    ("split_threshold", "sparse_oracle.cpp", "if (*ia < 0 || *ib < 0 || lo >= -0.3f) {",
     "if (*ia < 0 || *ib < 0 || lo >= -1.5f) {", SP_PINNED, "test_thin_sheet_in_closed_form"),
    ("mass_point", "sparse_oracle.cpp", "c = mul(1.0f / (float)cnt, c);",
     "c = mul(1.0f / (float)(cnt + 1), c);", SP_PINNED, "test_thin_sheet_in_closed_form"),
    ("face_orientation", "sparse_oracle.cpp",
     "if (dot(gn, gn) > 0.0f && dot(qn, qn) > 0.0f && dot(qn, gn) < 0.0f) {",
     "if (dot(gn, gn) > 0.0f && dot(qn, qn) > 0.0f && dot(qn, gn) > 0.0f) {", SP_PINNED,
     "test_smooth_topology_equals_the_dense_oracle"),
    ("face_cycle", "sparse_oracle.cpp",
     "const int t[4][3] = {{i, j - 1, k - 1}, {i, j, k - 1}, {i, j, k}, {i, j - 1, k}};",
     "const int t[4][3] = {{i, j - 1, k - 1}, {i, j, k}, {i, j, k - 1}, {i, j - 1, k}};",
     SP_PINNED, "test_smooth_topology_equals_the_dense_oracle"),
]


def _build(tmp_path, name, source=None, old=None, new=None):
    d = tmp_path / name
    d.mkdir(exist_ok=True)
    for src in SOURCES:
        text = open(os.path.join(ROOT, "oracle", src)).read()
        if src == source:
            assert text.count(old) == 1, (name, text.count(old))
            text = text.replace(old, new)
        (d / src).write_text(text)
    lib = str(d / "liboracle_dc.so")
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math",
                        "-fopenmp", "-fPIC", "-shared", "-o", lib,
                        *(str(d / src) for src in SOURCES)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return lib


def _run_pinned(lib, target):
    env = dict(os.environ, ORACLE_LIB=lib)
    return subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                           target], cwd=ROOT, env=env, capture_output=True, text=True,
                          timeout=300)


@pytest.fixture(scope="module")
def libs(tmp_path_factory):
    """Every library (the control and each mutant), built once, several compilers at a time."""
    from concurrent.futures import ThreadPoolExecutor
    root = tmp_path_factory.mktemp("mutants")
    jobs = [("control", None, None, None)] + [(m[0], m[1], m[2], m[3]) for m in MUTATIONS]
    with ThreadPoolExecutor(max_workers=max(1, min(8, os.cpu_count() or 1))) as ex:
        built = list(ex.map(lambda j: _build(root, *j), jobs))
    return dict(zip((j[0] for j in jobs), built))


def test_unmodified_copy_passes(libs):
    for target in (PINNED, POST_PINNED, M2S_PINNED, SP_PINNED):
        r = _run_pinned(libs["control"], target)
        assert r.returncode == 0, r.stdout[-2000:]


@pytest.mark.parametrize("name,source,old,new,path,test", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_mutation_is_caught(libs, name, source, old, new, path, test):
    r = _run_pinned(libs[name], f"{path}::{test}")
    assert r.returncode != 0, f"mutation {name} survived {test}:\n{r.stdout[-1500:]}"
