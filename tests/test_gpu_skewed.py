"""GPU parity on strongly skewed cells and ragged grids.

The tiled kernel drops candidate records by bounds on the distance from an
atom to a cell's or a wave's parallelepiped of points (its support function
along the cell vectors), and splits tiles into wave quadrants whose boxes
shrink at the grid's far faces.  These cases stress exactly that: cell
angles far from 90 degrees (so the boxes are far from spheres), grids with a
2-point axis, prime point counts (partial tiles and quadrants on every
face), and tile shapes whose wave quadrants are long and thin (16 x 2 x 4
points) or flat (2 x 8 x 8).

Tolerances as everywhere: FP64 grid 1e-9 relative to the oracle, FP32 grid
1e-3 K where E < 1e4 K (1e-6 relative above), in-cutoff counts exact,
Widom energies 1e-3 K where E < 1e4 K, K_H 1e-4 relative.
"""
import numpy as np
import pytest

from emptyroom_amd import api, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

CELLS = [
    (12.0, 8.0, 10.0, 120.0, 62.0, 118.0),
    (7.5, 13.0, 9.0, 100.0, 70.0, 130.0),
    (11.0, 9.0, 10.0, 58.0, 62.0, 64.0),
]
GRIDS = [[13, 19, 23], [2, 17, 9], [31, 5, 29]]


@pytest.fixture(scope="module")
def sctx():
    c = api.Context(0)
    yield c
    c.close()


def set_tile(c, tile):
    t = c.tuning
    t.tile[0], t.tile[1], t.tile[2] = tile
    c.tuning = t


def check_fp32(g32, ref):
    mask = np.abs(ref) < 1e4
    if mask.any():
        assert np.max(np.abs(g32.astype(np.float64)[mask] - ref[mask])) <= 1e-3
    hi = ~mask & (np.abs(ref) < 3.0e38)
    if hi.any():
        rel = np.abs(g32.astype(np.float64)[hi] - ref[hi]) / np.abs(ref[hi])
        assert rel.max() < 1e-6, rel.max()


@pytest.mark.parametrize("tile", [(8, 8, 8), (16, 4, 8), (2, 16, 16)])
@pytest.mark.parametrize("ci", range(len(CELLS)))
def test_skewed_grids_match_oracle(sctx, ci, tile):
    print(f"skewed cell {CELLS[ci]} tile {tile}", flush=True)
    default = list(sctx.tuning.tile)
    set_tile(sctx, tile)
    try:
        f = synth.small_framework(31 + ci, n_atoms=12, cell=CELLS[ci])  # 8-27% of points < 1e4 K
        a = f.particles()
        fw = api.Framework(sctx, a[:, :3], a[:, 3], a[:, 4], f.t, f.rep, f.rc2)
        for n in GRIDS:
            ref = oracle.lj_grid(f.t, a, f.rep, f.rc2, n)
            g64 = api.compute_grid(sctx, fw, n, api.ER_FP64)
            rel = np.abs(g64 - ref) / np.maximum(1.0, np.abs(ref))
            assert rel.max() < 1e-9, (n, rel.max())
            check_fp32(api.compute_grid(sctx, fw, n, api.ER_FP32), ref)
            assert api.count_incut_grid(sctx, fw, n) == \
                oracle.count_incut_grid(f.t, a, f.rep, f.rc2, n)
        fw.close()
    finally:
        set_tile(sctx, default)


@pytest.mark.parametrize("ci", range(len(CELLS)))
def test_skewed_widom_matches_oracle(sctx, ci):
    print(f"skewed widom cell {CELLS[ci]}", flush=True)
    f = synth.small_framework(41 + ci, n_atoms=24, cell=CELLS[ci])
    a = f.particles()
    fw = api.Framework(sctx, a[:, :3], a[:, 3], a[:, 4], f.t, f.rep, f.rc2)
    nins, seed, T = 12000, 77 + ci, 310.0
    sb, seb, n, e = api.widom(sctx, fw, nins, T, seed, energies=True)
    rb, reb, re = oracle.widom(f.t, a, f.rep, f.rc2, T, seed, 0, nins, energies=True)
    assert n == nins
    mask = re < 1e4
    assert mask.sum() > 50
    assert np.max(np.abs(e[mask] - re[mask])) <= 1e-3
    assert abs(sb - rb) / rb <= 1e-4
    sb2, _, _ = api.widom(sctx, fw, nins, T, seed)
    assert abs(sb2 - rb) / rb <= 1e-4
    fw.close()
