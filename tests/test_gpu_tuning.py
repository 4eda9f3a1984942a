"""GPU parity of every kernel shape the tuning knobs select.

A lane evaluates 1, 2 or 4 grid points (tile = 256 x points-per-lane points).
With several points per lane, the points of a lane differ only in i when
tile_y * tile_z divides 256, and on a cell whose first vector lies along x
(what cell parameters give) they share y and z, so the kernel forms the y and
z terms of the expanded r^2 once per record.  These tests cover that shared
form, the per-point form (a rotated cell matrix, and a tile whose y-z face
does not divide 256), partial tiles on every face (odd point counts), and
Widom bins, which take at most 2 insertions per lane.

Tolerances as in the other parity tests: FP64 grid 1e-9 relative, FP32 grid
1e-3 K absolute where E < 1e4 K, counts and hard-sphere grids exact, Widom
energies 1e-3 K.  Widom energies do not depend on points per lane (the bin
layout and the LDS list are the same), so they must be bit-identical.
"""
import numpy as np
import pytest

from emptyroom_amd import api, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

# (points per lane, tile)
SHAPES = [
    (1, (8, 8, 4)),
    (2, (8, 8, 8)),
    (2, (16, 4, 8)),
    (4, (16, 8, 8)),
    (4, (8, 8, 16)),
    (4, (4, 16, 16)),   # y-z face of 256: points of a lane one i apart
    (4, (2, 32, 16)),   # y-z face of 512: points of a lane differ in j
]

CELL = (10.5, 11.0, 12.0, 80.0, 95.0, 105.0)


def rotation(seed):
    q, r = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


@pytest.fixture(scope="module")
def tctx():
    c = api.Context(0)
    yield c
    c.close()


def set_shape(c, ppl, tile, region_len=None):
    t = c.tuning
    t.points_per_lane = ppl
    t.tile[0], t.tile[1], t.tile[2] = tile
    if region_len is not None:
        t.region_len = region_len
    c.tuning = t


def check_fp32(g32, ref):
    mask = np.abs(ref) < 1e4
    err = float(np.max(np.abs(g32.astype(np.float64)[mask] - ref[mask])))
    assert err <= 1e-3, f"max abs error {err} K over |E| < 1e4 K"
    hi = ~mask & (np.abs(ref) < 3.0e38)
    if hi.any():
        rel = np.abs(g32.astype(np.float64)[hi] - ref[hi]) / np.abs(ref[hi])
        assert rel.max() < 1e-6, rel.max()


@pytest.fixture(scope="module")
def frameworks():
    f = synth.small_framework(71, n_atoms=60, cell=CELL)
    a = f.particles()
    rot = rotation(5)
    return [("upper", f, f.t, a), ("rotated", f, rot @ f.t, a)]


@pytest.fixture(scope="module")
def references(frameworks):
    n = [29, 23, 31]
    out = {}
    for name, f, t, a in frameworks:
        out[name] = (oracle.lj_grid(t, a, f.rep, f.rc2, n),
                     oracle.count_incut_grid(t, a, f.rep, f.rc2, n))
    return n, out


@pytest.mark.parametrize("ppl,tile", SHAPES)
def test_grid_shapes_match_oracle(tctx, frameworks, references, ppl, tile):
    n, refs = references
    for region_len in (0.0, 3.0):
        print(f"tuning ppl={ppl} tile={tile} region_len={region_len}", flush=True)
        set_shape(tctx, ppl, tile, region_len)
        for name, f, t, a in frameworks:
            ref, cnt = refs[name]
            fw = api.Framework(tctx, a[:, :3], a[:, 3], a[:, 4], t, f.rep, f.rc2)
            g64 = api.compute_grid(tctx, fw, n, api.ER_FP64)
            rel = np.abs(g64 - ref) / np.maximum(1.0, np.abs(ref))
            assert rel.max() < 1e-9, (name, rel.max())
            check_fp32(api.compute_grid(tctx, fw, n, api.ER_FP32), ref)
            assert api.count_incut_grid(tctx, fw, n) == cnt, name
            fw.close()
    tctx.tuning = api.default_tuning()


@pytest.mark.parametrize("ppl,tile", [SHAPES[0], SHAPES[1], SHAPES[3], SHAPES[6]])
def test_accessibility_shapes_match_oracle(tctx, frameworks, ppl, tile):
    print(f"accessibility ppl={ppl} tile={tile}", flush=True)
    set_shape(tctx, ppl, tile)
    for name, f, t, a in frameworks:
        n = [23, 21, 25]
        got = api.accessibility_grid(tctx, a[:, :3], a[:, 3], a[:, 4], t, n)
        ref = oracle.accessibility_grid(t, a, n)
        assert np.array_equal(got, ref), name
    tctx.tuning = api.default_tuning()


def test_widom_energies_independent_of_points_per_lane(tctx, frameworks):
    name, f, t, a = frameworks[1]
    nins, seed, T = 30000, 17, 298.0
    rb, reb, re = oracle.widom(t, a, f.rep, f.rc2, T, seed, 0, nins, energies=True)
    first = None
    for ppl, tile in SHAPES[:4]:
        print(f"widom ppl={ppl}", flush=True)
        set_shape(tctx, ppl, tile)
        fw = api.Framework(tctx, a[:, :3], a[:, 3], a[:, 4], t, f.rep, f.rc2)
        sb, seb, n, e = api.widom(tctx, fw, nins, T, seed, energies=True)
        fw.close()
        assert n == nins
        mask = re < 1e4
        assert np.max(np.abs(e[mask] - re[mask])) <= 1e-3
        assert abs(sb - rb) / rb <= 1e-4
        if first is None:
            first = (e.copy(), sb, seb)
        else:
            assert np.array_equal(e, first[0])
            assert sb == first[1] and seb == first[2]
    tctx.tuning = api.default_tuning()
