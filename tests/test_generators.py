"""The dc_cli generators (sphere, box, plane on [-1, 1]^3; the reference's src/cpp/dc_cli.cpp:49-118)
against their formulas evaluated in FP32 with NumPy, bit for bit, and the .sdf writer/reader
round trip (the layout of dc_cli.cpp:31-47 and scripts/test_glb_benchmark.py:20-25). CPU only."""
import numpy as np
import pytest

from contouring_hip import io, synth


def formula(kind, res):
    h = np.float32(2.0) / np.float32(res - 1)
    ax = np.float32(-1.0) + np.arange(res, dtype=np.float32) * h  # ox + i * vx in FP32
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    if kind == "sphere":
        return np.sqrt((x * x + y * y) + z * z) - np.float32(0.5)
    if kind == "box":
        dx, dy, dz = (np.abs(t) - np.float32(0.5) for t in (x, y, z))
        return np.maximum(dx, np.maximum(dy, dz))
    return z.copy()


@pytest.mark.parametrize("kind", ["sphere", "box", "plane"])
@pytest.mark.parametrize("res", [2, 17, 40, 64])
def test_generator_matches_formula(kind, res):
    g = synth.cli_generate(kind, res)
    want = formula(kind, res).ravel()
    np.testing.assert_array_equal(np.asarray(g.values, np.float32), want)
    assert np.float32(g.vx) == np.float32(2.0) / np.float32(res - 1)
    assert (g.ox, g.oy, g.oz) == (-1.0, -1.0, -1.0)


def test_sdf_round_trip(tmp_path):
    g = synth.cli_generate("box", 20)
    p = tmp_path / "g.sdf"
    io.save_sdf(p, g)
    raw = p.read_bytes()
    assert len(raw) == 36 + 4 * 20 ** 3
    assert np.frombuffer(raw[:12], np.int32).tolist() == [20, 20, 20]
    h = io.load_sdf(p)
    assert (h.nx, h.ny, h.nz) == (20, 20, 20)
    assert (np.float32(h.vx), h.ox) == (np.float32(g.vx), -1.0)
    np.testing.assert_array_equal(h.values, np.asarray(g.values, np.float32))
    (tmp_path / "short.sdf").write_bytes(raw[:-4])
    with pytest.raises(ValueError):
        io.load_sdf(tmp_path / "short.sdf")
    bad = bytearray(raw)
    bad[0:4] = np.array([1], np.int32).tobytes()
    (tmp_path / "bad.sdf").write_bytes(bytes(bad))
    with pytest.raises(ValueError):
        io.load_sdf(tmp_path / "bad.sdf")


def test_obj_round_trip(tmp_path):
    V = np.array([[0, 0, 0], [1, 0.5, 0], [1, 1, 0.25], [0, 1, 0]], np.float32)
    F = np.array([[0, 1, 2, 3], [3, 2, 1, 0]], np.int32)
    p = tmp_path / "m.obj"
    io.save_obj(p, V.ravel(), F.ravel())
    text = p.read_text().splitlines()
    assert text[4] == "f 1 2 3 4"
    V2, F2 = io.load_obj(p)
    np.testing.assert_array_equal(V2, V)
    np.testing.assert_array_equal(F2, F)
