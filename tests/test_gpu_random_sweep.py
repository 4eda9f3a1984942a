"""Seeded random sweep of the tiled kernel against the oracle.

Each case draws a framework (5-60 atoms, cell edges 7-16 A, angles 60-120
degrees), grid dimensions (2-37 points per axis), a tile shape holding 512
points (the library's two points per lane), the Widom insertions per bin and
a temperature, then checks: the FP64 grid (1e-9 relative), the FP32 grid
(1e-3 K where E < 1e4 K, 1e-6 relative above), the in-cutoff count (exact),
Widom energies (1e-3 K where E < 1e4 K) and K_H with the saturation stop on
(1e-4 relative) against the oracle.  Cases are kept small enough for the
oracle to finish in well under a second each.
"""
import math

import numpy as np
import pytest

from emptyroom_amd import api, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

TILES = [(8, 8, 8), (16, 4, 8), (4, 8, 16), (2, 16, 16), (16, 16, 2), (8, 4, 16)]
N_CASES = 24


def draw(seed):
    rng = np.random.default_rng(seed)
    while True:
        cell = tuple(rng.uniform(7.0, 16.0, 3)) + tuple(rng.uniform(60.0, 120.0, 3))
        ca, cb, cg = (math.cos(math.radians(x)) for x in cell[3:])
        if 1 - ca * ca - cb * cb - cg * cg + 2 * ca * cb * cg > 0.05:
            break
    n_atoms = int(rng.integers(5, 61))
    grid = [int(x) for x in rng.integers(2, 38, 3)]
    tile = TILES[int(rng.integers(0, len(TILES)))]
    per_bin = int(rng.choice([64, 240, 480, 900]))
    temperature = float(rng.uniform(250.0, 350.0))
    return cell, n_atoms, grid, tile, per_bin, temperature


@pytest.fixture(scope="module")
def rctx():
    c = api.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("seed", range(N_CASES))
def test_random_case_matches_oracle(rctx, seed):
    cell, n_atoms, grid, tile, per_bin, temperature = draw(1000 + seed)
    print(f"random case {seed}: cell {np.round(cell, 2).tolist()} atoms {n_atoms} grid {grid} "
          f"tile {tile} per_bin {per_bin}", flush=True)
    t = rctx.tuning
    saved = (list(t.tile), t.widom_per_bin)
    t.tile[0], t.tile[1], t.tile[2] = tile
    t.widom_per_bin = per_bin
    rctx.tuning = t
    try:
        f = synth.small_framework(2000 + seed, n_atoms=n_atoms, cell=cell)
        a = f.particles()
        fw = api.Framework(rctx, a[:, :3], a[:, 3], a[:, 4], f.t, f.rep, f.rc2)
        ref = oracle.lj_grid(f.t, a, f.rep, f.rc2, grid)
        g64 = api.compute_grid(rctx, fw, grid, api.ER_FP64)
        assert (np.abs(g64 - ref) / np.maximum(1.0, np.abs(ref))).max() < 1e-9
        g32 = api.compute_grid(rctx, fw, grid, api.ER_FP32).astype(np.float64)
        lo = np.abs(ref) < 1e4
        if lo.any():
            assert np.max(np.abs(g32[lo] - ref[lo])) <= 1e-3
        hi = ~lo & (np.abs(ref) < 3.0e38)
        if hi.any():
            assert (np.abs(g32[hi] - ref[hi]) / np.abs(ref[hi])).max() < 1e-6
        assert api.count_incut_grid(rctx, fw, grid) == \
            oracle.count_incut_grid(f.t, a, f.rep, f.rc2, grid)
        nins, wseed = 6000, 77 + seed
        sb, seb, n, e = api.widom(rctx, fw, nins, temperature, wseed, energies=True)
        rb, reb, re = oracle.widom(f.t, a, f.rep, f.rc2, temperature, wseed, 0, nins,
                                   energies=True)
        assert n == nins
        m = re < 1e4
        if m.any():
            assert np.max(np.abs(e[m] - re[m])) <= 1e-3
        sb2, _, _ = api.widom(rctx, fw, nins, temperature, wseed)
        if rb > 0:
            assert abs(sb - rb) / rb <= 1e-4
            assert abs(sb2 - rb) / rb <= 1e-4
        else:
            assert sb == 0.0 and sb2 == 0.0
        fw.close()
    finally:
        t = rctx.tuning
        t.tile[0], t.tile[1], t.tile[2] = saved[0]
        t.widom_per_bin = saved[1]
        rctx.tuning = t
