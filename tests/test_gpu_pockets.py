"""GPU pocket blocking and grid-interpolated Widom vs the oracle (bit-exact
segment ids and blocked grids; K_H at the same insertion points)."""
import numpy as np
import pytest

from emptyroom_amd import api, synth
from oracle import oracle

pytestmark = pytest.mark.gpu
T = 298.0


def random_grid(seed, n, fill):
    r = synth.Rng(seed).uniform(int(np.prod(n))).reshape(n)
    # accessible voxels below 15 T, inaccessible above; distinct values
    return np.where(r < fill, 100.0 * r, 1e5 + r).astype(np.float32)


def check_against_oracle(ctx, g32):
    ref_n, ref_blocked, ref_seg = oracle.find_and_block_pockets(g32.astype(np.float64), T)
    grid = api.Grid.upload(ctx, g32)
    n, seg = api.pockets_label(ctx, grid, T, block=True, segments=True)
    blocked = grid.download()
    grid.close()
    assert n == ref_n
    assert np.array_equal(seg, ref_seg)
    assert np.array_equal(blocked, ref_blocked.astype(np.float32))
    return n


@pytest.mark.parametrize("seed,n,fill", [(1, (20, 18, 16), 0.25), (2, (24, 24, 24), 0.31),
                                         (3, (17, 23, 19), 0.40), (4, (40, 12, 30), 0.33),
                                         (5, (9, 9, 9), 0.60), (6, (33, 35, 31), 0.29)])
def test_random_grids_bit_exact(ctx, seed, n, fill):
    print(f"pockets random seed={seed} n={n} fill={fill}", flush=True)
    check_against_oracle(ctx, random_grid(seed, n, fill))


def test_hand_built_cases(ctx):
    wall = np.float32(1e5)
    g = np.full((8, 8, 8), wall, np.float32)
    g[3:5, 3:5, 3:5] = 0.0
    assert check_against_oracle(ctx, g) == 1
    g = np.full((8, 7, 6), wall, np.float32)
    g[:, 3, 2] = 0.0
    assert check_against_oracle(ctx, g) == 0
    g = np.full((6, 6, 3), wall, np.float32)
    g[0, 0, 1] = 0.0
    for x, y in [(5, 0), (5, 1), (4, 1), (4, 2), (3, 2), (3, 3), (2, 3), (2, 4), (1, 4), (1, 5),
                 (0, 5)]:
        g[x, y, 1] = 0.0
    assert check_against_oracle(ctx, g) == 0
    g[0, 5, 1] = wall
    assert check_against_oracle(ctx, g) == 2
    g = np.full((5, 5, 5), wall, np.float32)
    assert check_against_oracle(ctx, g) == 0  # no accessible voxel at all


def framework_grid(ctx, seed):
    f = synth.small_framework(seed, n_atoms=70, cell=(15.0, 16.0, 14.0, 85.0, 95.0, 100.0))
    a = f.particles()
    fw = api.Framework(ctx, a[:, :3], a[:, 3], a[:, 4], f.t, f.rep, f.rc2)
    n = [int(np.ceil(L / 0.5)) for L in f.cell[:3]]
    g = api.compute_grid(ctx, fw, n, api.ER_FP32)
    fw.close()
    return g


@pytest.mark.parametrize("seed", [21, 22])
def test_framework_grid_pockets_and_grid_widom(ctx, seed):
    print(f"pockets + grid Widom on a framework grid seed={seed}", flush=True)
    g32 = framework_grid(ctx, seed)
    assert (g32 < 15 * T).mean() > 0.05  # accessible space exists
    check_against_oracle(ctx, g32)
    grid = api.Grid.upload(ctx, g32)
    api.pockets_label(ctx, grid, T, block=True)
    blocked = grid.download()
    nins, s = 50000, 77
    sb, seb, n, e = api.widom_grid(ctx, grid, nins, T, s, energies=True)
    rb, reb, re = oracle.widom_grid(blocked.astype(np.float64), T, s, 0, nins, energies=True)
    grid.close()
    assert n == nins
    assert np.allclose(e, re, rtol=1e-12, atol=1e-9)
    assert abs(sb - rb) / rb < 1e-9
    assert abs(seb / sb - reb / rb) < 1e-6 * max(1.0, abs(reb / rb))


def test_fp64_grid_is_refused(ctx):
    f = synth.small_framework(3, n_atoms=10)
    a = f.particles()
    fw = api.Framework(ctx, a[:, :3], a[:, 3], a[:, 4], f.t, f.rep, f.rc2)
    g = api.Grid.compute(ctx, fw, [5, 5, 5], api.ER_FP64)
    with pytest.raises(api.EmptyRoomError):
        api.pockets_label(ctx, g, T)
    g.close()
    fw.close()


def test_batch_matches_oracle_per_grid(ctx):
    """Many grids of different shapes labelled in one call: every grid's pocket
    count and blocked energies equal the oracle's on that grid alone."""
    shapes = [((20, 18, 16), 0.25), ((9, 9, 9), 0.6), ((33, 35, 31), 0.29), ((17, 23, 19), 0.4),
              ((40, 12, 30), 0.33), ((8, 8, 8), 0.0), ((2, 3, 2), 1.0)]
    grids32 = [random_grid(100 + i, n, fill) for i, (n, fill) in enumerate(shapes)]
    print(f"pockets batch of {len(grids32)} grids", flush=True)
    handles = [api.Grid.upload(ctx, g) for g in grids32]
    counts = api.pockets_label_batch(ctx, handles, T, block=True)
    for g32, h, n in zip(grids32, handles, counts):
        ref_n, ref_blocked, _ = oracle.find_and_block_pockets(g32.astype(np.float64), T)
        assert n == ref_n
        assert np.array_equal(h.download(), ref_blocked.astype(np.float32))
        h.close()


def test_pockets_on_loaded_grid_file(ctx, tmp_path):
    """A text grid file (the reference's txt format) loaded as a resident
    FP32 grid: pockets and blocked energies as the oracle on the same values."""
    print("pockets on a loaded txt grid", flush=True)
    g = random_grid(31, (21, 19, 17), 0.35).astype(np.float64)
    path = tmp_path / "g.txt"
    api.write_grid_file(str(path), "txt", g, np.diag([10.0, 9.0, 8.0]), threads=2)
    values = api.read_grid_file(str(path)).astype(np.float32)
    grid = api.Grid.load(ctx, str(path))
    assert np.array_equal(grid.download(), values)
    n, _ = api.pockets_label(ctx, grid, T, block=True)
    ref_n, ref_blocked, _ = oracle.find_and_block_pockets(values.astype(np.float64), T)
    assert n == ref_n
    assert np.array_equal(grid.download(), ref_blocked.astype(np.float32))
    grid.close()
