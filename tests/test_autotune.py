"""The autotuner's search and the host-side tuning helpers, on the CPU.

The search is checked against cost functions whose optimum is known (a
separable bowl, a coupled one where the best class tile depends on points
per lane, a noisy plateau that must not move the incumbent); the size-class
rule against grids of known spacing; the tuning file and struct round trips;
and the library's rejection of class tiles it cannot run.
"""
import numpy as np
import pytest

from emptyroom_amd import api, autotune


def test_tiles_for_shapes_are_valid():
    for ppl in (1, 2, 4):
        tiles = autotune.tiles_for(ppl)
        assert tiles and len(set(tiles)) == len(tiles)
        for a, b, c in tiles:
            assert a * b * c == 256 * ppl
            assert 256 % (b * c) == 0
        assert (8, 8, 8) in autotune.tiles_for(2)


def start_cfg():
    return api.default_tuning().to_dict()


def test_search_finds_separable_optimum():
    want = dict(points_per_lane=2, lds_records=1280, widom_per_bin=720, region_len=2.5)

    def cost(cfg):
        c = 1.0
        for k, v in want.items():
            c += 0.0 if cfg[k] == v else 0.1 + 0.001 * abs(float(cfg[k]) - float(v))
        return c

    best, t, hist = autotune.coordinate_search(cost, start_cfg(), autotune.knobs([0]), rounds=3)
    for k, v in want.items():
        assert best[k] == v
    assert t == 1.0
    # memoised: every configuration measured once
    keys = [autotune._key(c) for c, _ in hist]
    assert len(keys) == len(set(keys))


def test_class_tile_search_follows_points_per_lane():
    # best: 2 points per lane, class 1 uses (16, 8, 4), class 0 the default
    def cost(cfg):
        c = 1.0 if cfg["points_per_lane"] == 2 else 1.5
        if cfg["class_tile"][1] == [16, 8, 4]:
            c -= 0.2
        if cfg["class_tile"][0] != [0, 0, 0]:
            c += 0.05
        return c

    best, t, _ = autotune.coordinate_search(cost, start_cfg(), autotune.knobs([0, 1]), rounds=3)
    assert best["points_per_lane"] == 2 and best["tile"] == [8, 8, 8]
    assert best["class_tile"][1] == [16, 8, 4]
    assert best["class_tile"][0] == [0, 0, 0]
    api.Tuning.from_dict(best)  # a valid struct
    assert abs(t - 0.8) < 1e-12


def test_noise_within_margin_keeps_incumbent():
    rng = np.random.default_rng(0)
    start = start_cfg()

    def cost(cfg):
        return 1.0 + (0.0 if cfg == start else rng.uniform(-0.004, 0.004))

    best, _, _ = autotune.coordinate_search(cost, start, autotune.knobs([0, 1, 2]), margin=0.01)
    assert best == start


def test_size_classes_by_spacing():
    tu = api.default_tuning()
    t = np.diag([20.0, 20.0, 20.0])
    assert api.grid_tile_for(tu, t, [201, 201, 201])[1] == 0   # 0.1 A
    assert api.grid_tile_for(tu, t, [101, 101, 101])[1] == 1   # 0.2 A
    assert api.grid_tile_for(tu, t, [41, 41, 41])[1] == 2      # 0.5 A
    # geometric mean of the three spacings: 0.1 * 0.2 * 0.4 -> 0.2 A
    assert api.grid_tile_for(tu, t, [201, 101, 51])[1] == 1
    d = tu.to_dict()
    tile = [16, 8, 2 * d["points_per_lane"]]
    d["class_tile"] = [[0, 0, 0], tile, [0, 0, 0]]
    tu2 = api.Tuning.from_dict(d)
    assert api.grid_tile_for(tu2, t, [101, 101, 101]) == (tuple(tile), 1)
    assert api.grid_tile_for(tu2, t, [201, 201, 201]) == (tuple(d["tile"]), 0)


def test_invalid_class_tile_rejected():
    d = api.default_tuning().to_dict()
    d["class_tile"] = [[4, 4, 4], [0, 0, 0], [0, 0, 0]]
    with pytest.raises(api.EmptyRoomError):
        api.grid_tile_for(api.Tuning.from_dict(d), np.eye(3) * 10, [11, 11, 11])
    d = api.default_tuning().to_dict()
    d["class_spacing"] = [0.3, 0.2]
    with pytest.raises(api.EmptyRoomError):
        api.grid_tile_for(api.Tuning.from_dict(d), np.eye(3) * 10, [11, 11, 11])


def test_tuning_file_round_trip(tmp_path):
    d = api.default_tuning().to_dict()
    d.update(points_per_lane=2, tile=[8, 8, 8], lds_records=1280,
             class_tile=[[0, 0, 0], [16, 8, 4], [0, 0, 0]])
    path = tmp_path / "t.json"
    autotune.save(str(path), {"tuning": d, "history": []}, meta={"device": "test"})
    assert autotune.load(str(path)).to_dict() == d
    with pytest.raises(KeyError):
        api.Tuning.from_dict({"no_such_knob": 1})


def test_tuning_struct_matches_library_layout():
    """The library writes exactly sizeof(er_tuning) bytes: a canary past the
    Python struct must survive (a mismatch corrupts the caller's heap)."""
    import ctypes
    n = ctypes.sizeof(api.Tuning)
    buf = (ctypes.c_ubyte * (n + 64))(*([0xA5] * (n + 64)))
    assert api.lib().er_tuning_defaults(ctypes.cast(buf, ctypes.POINTER(api.Tuning))) == 0
    assert all(b == 0xA5 for b in bytes(buf)[n:])
    t = api.Tuning.from_buffer_copy(bytes(buf)[:n])
    assert t.to_dict() == api.default_tuning().to_dict()
    assert t.lds_records >= 512 and t.class_spacing[0] > 0
