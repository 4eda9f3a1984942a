"""Grid files on the CPU: the text reader (src/gcmc.cc:286-346 semantics)
against the text writer (whose bytes the layout test pins to
src/writegrid.cu:260-269) and against numpy's own parse of the same text;
the binary side-format round trip; malformed files are errors."""
import numpy as np
import pytest

from emptyroom_amd import api


def sample_grid(shape, seed=0):
    rng = np.random.default_rng(seed)
    g = rng.normal(0, 300, size=shape)
    g.flat[::7] = 1e13          # blocked voxels
    g.flat[3::11] = -2.5e-7     # small magnitudes
    g.flat[5::13] = 4.47e3      # near the 15 T threshold
    return g


def test_txt_round_trip_matches_text_values(tmp_path):
    g = sample_grid((7, 5, 9))
    t = np.diag([10.0, 11.0, 12.0])
    path = tmp_path / "g.txt"
    api.write_grid_file(str(path), "txt", g, t, threads=3)
    got = api.read_grid_file(str(path))
    lines = path.read_text().split("\n")
    assert lines[0].startswith("7 5 9 ")
    ref = np.array([float(v) for line in lines[1:] for v in line.split()]).reshape(7, 5, 9)
    assert np.array_equal(got, ref)
    assert np.array_equal(got, np.array([float(f"{v:13.6E}") for v in g.flat]).reshape(g.shape))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_binary_round_trip(tmp_path, dtype):
    g = sample_grid((6, 4, 5), seed=1).astype(dtype)
    t = np.array([[10.0, 0.3, 0.2], [0.0, 9.0, 0.1], [0.0, 0.0, 8.0]])
    path = tmp_path / "g.grid"
    api.write_grid_binary(str(path), g, t)
    raw = path.read_bytes()
    assert raw[:8] == b"ERGRID01" and len(raw) == 128 + g.size * g.itemsize
    hdr = np.frombuffer(raw[8:32], dtype=np.int32)
    assert list(hdr) == [1, 1 if dtype == np.float64 else 0, 6, 4, 5, 0]
    assert np.array_equal(np.frombuffer(raw[32:104], dtype=np.float64).reshape(3, 3), t)
    assert np.array_equal(api.read_grid_file(str(path)), g.astype(np.float64))


def test_malformed_text_grids_are_errors(tmp_path):
    g = sample_grid((3, 2, 4), seed=2)
    path = tmp_path / "g.txt"
    api.write_grid_file(str(path), "txt", g, np.eye(3) * 5, threads=1)
    text = path.read_text().split("\n")
    short = tmp_path / "short.txt"
    short.write_text("\n".join(text[:-3]) + "\n")  # one z-pencil missing
    with pytest.raises(api.EmptyRoomError):
        api.read_grid_file(str(short))
    cut = tmp_path / "cut.txt"
    bad = list(text)
    bad[2] = " ".join(bad[2].split()[:3])  # a short line
    cut.write_text("\n".join(bad))
    with pytest.raises(api.EmptyRoomError):
        api.read_grid_file(str(cut))
    with pytest.raises(api.EmptyRoomError):
        api.read_grid_file(str(tmp_path / "missing.txt"))
    hdr = tmp_path / "hdr.txt"
    hdr.write_text("x y z\n")
    with pytest.raises(api.EmptyRoomError):
        api.read_grid_file(str(hdr))
