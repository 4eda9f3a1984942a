"""GPU parity of the LJ energy grid and Widom insertion against the oracle.

Tolerances (from the workload definition): FP32 grid max abs error <= 1e-3 K
over points with E < 1e4 K; K_H relative error <= 1e-4 evaluated at the same
insertion points.  The FP64 grid runs the reference arithmetic on every pair
and must agree to 1e-9 relative.  Integer work (in-cutoff counts) is exact.
"""
import numpy as np
import pytest

from emptyroom_amd import api, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

GRID_TOL_K = 1e-3
KH_TOL = 1e-4

CELLS = [
    (10.0, 11.0, 12.0, 85.0, 95.0, 100.0),
    (9.0, 14.0, 8.5, 70.0, 110.0, 65.0),
    (16.0, 16.0, 16.0, 90.0, 90.0, 90.0),
]


def upload(ctx, f):
    a = f.particles()
    return api.Framework(ctx, a[:, :3], a[:, 3], a[:, 4], f.t, f.rep, f.rc2), a


def check_fp32(g32, ref):
    mask = ref < 1e4
    err = float(np.max(np.abs(g32.astype(np.float64)[mask] - ref[mask]))) if mask.any() else 0.0
    assert err <= GRID_TOL_K, f"max abs error {err} K over E < 1e4 K"
    hi = ~mask & (ref < 3.0e38)
    if hi.any():
        rel = np.abs(g32.astype(np.float64)[hi] - ref[hi]) / ref[hi]
        assert rel.max() < 1e-6
    return err


@pytest.mark.parametrize("cell", CELLS)
@pytest.mark.parametrize("seed", [1, 2])
def test_grid_matches_oracle(ctx, cell, seed):
    print(f"grid parity cell={cell} seed={seed}", flush=True)
    f = synth.small_framework(seed, n_atoms=60, cell=cell)
    fw, atoms = upload(ctx, f)
    n = [13, 11, 12]
    ref = oracle.lj_grid(f.t, atoms, f.rep, f.rc2, n)
    g64 = api.compute_grid(ctx, fw, n, api.ER_FP64)
    rel = np.abs(g64 - ref) / np.maximum(1.0, np.abs(ref))
    assert rel.max() < 1e-9, rel.max()
    g32 = api.compute_grid(ctx, fw, n, api.ER_FP32)
    check_fp32(g32, ref)
    fw.close()


def test_grid_count_incut_exact(ctx):
    f = synth.small_framework(3, n_atoms=50, cell=CELLS[1])
    fw, atoms = upload(ctx, f)
    n = [17, 9, 14]
    assert api.count_incut_grid(ctx, fw, n) == oracle.count_incut_grid(f.t, atoms, f.rep,
                                                                        f.rc2, n)
    fw.close()


@pytest.mark.parametrize("cell", CELLS)
def test_widom_matches_oracle(ctx, cell):
    print(f"widom parity cell={cell}", flush=True)
    f = synth.small_framework(11, n_atoms=40, cell=cell)
    fw, atoms = upload(ctx, f)
    nins, seed, T = 20000, 99, 298.0
    sb, seb, n, e = api.widom(ctx, fw, nins, T, seed, energies=True)
    rb, reb, re = oracle.widom(f.t, atoms, f.rep, f.rc2, T, seed, 0, nins, energies=True)
    assert n == nins
    mask = re < 1e4
    assert np.max(np.abs(e[mask] - re[mask])) <= GRID_TOL_K
    assert abs(sb - rb) / rb <= KH_TOL
    assert abs(seb / sb - reb / rb) <= 1e-3 * max(1.0, abs(reb / rb))
    fw.close()


def test_config1_full(ctx):
    """The reference's CPU-sized case in full: 50^3 grid and 1e5 insertions."""
    print("config-1 full parity", flush=True)
    f = synth.config1()
    fw, atoms = upload(ctx, f)
    n = f.grid_n()
    assert n == [50, 50, 50]
    ref = oracle.lj_grid(f.t, atoms, f.rep, f.rc2, n)
    g32 = api.compute_grid(ctx, fw, n, api.ER_FP32)
    check_fp32(g32, ref)
    g64 = api.compute_grid(ctx, fw, n, api.ER_FP64)
    assert (np.abs(g64 - ref) / np.maximum(1.0, np.abs(ref))).max() < 1e-9
    sb, seb, _ = api.widom(ctx, fw, f.n_insertions, f.temperature, f.seed)
    rb, reb = oracle.widom(f.t, atoms, f.rep, f.rc2, f.temperature, f.seed, 0, f.n_insertions)
    assert abs(sb - rb) / rb <= KH_TOL
    fw.close()


def test_batch_bit_identical_to_plain(ctx):
    fs = [synth.small_framework(s, n_atoms=30 + 10 * s, cell=CELLS[s % 3]) for s in range(4)]
    fws = [upload(ctx, f)[0] for f in fs]
    tasks = []
    for i, (f, fw) in enumerate(zip(fs, fws)):
        tasks.append(dict(fw=fw, grid_n=[9 + i, 10, 11], precision=api.ER_FP32,
                          n_insertions=5000 + 1000 * i, temperature=250.0 + 20 * i, seed=7 + i))
    for use_graph in (False, True):
        b = api.Batch(ctx, tasks, use_graph=use_graph)
        for rep in range(2):
            b.run()
            b.wait()
            for i, t in enumerate(tasks):
                g = api.compute_grid(ctx, t["fw"], t["grid_n"], api.ER_FP32)
                assert np.array_equal(b.grid(i), g)
                sb, seb, _ = api.widom(ctx, t["fw"], t["n_insertions"], t["temperature"],
                                       t["seed"])
                wb = b.widom(i)
                assert wb[0] == sb and wb[1] == seb
        b.close()
    for fw in fws:
        fw.close()


def test_empty_framework_and_clamp(ctx):
    # no atoms: E == 0 everywhere and every Boltzmann factor is 1
    t = np.diag([10.0, 10.0, 10.0])
    fw = api.Framework(ctx, np.zeros((0, 3)), np.zeros(0), np.zeros(0), t, [1, 1, 1], 156.25)
    assert np.all(api.compute_grid(ctx, fw, [5, 5, 5], api.ER_FP32) == 0.0)
    sb, seb, n = api.widom(ctx, fw, 1000, 300.0, 1)
    assert sb == 1000.0 and seb == 0.0
    fw.close()
    # one atom exactly on grid point (0,0,0): the reference clamps r^2 at 1e-24
    frac = np.array([[0.0, 0.0, 0.0]])
    fw = api.Framework(ctx, frac, np.array([100.0]), np.array([9.0]), t, [1, 1, 1], 156.25)
    ref = oracle.lj_grid(t, np.array([[0, 0, 0, 100.0, 9.0]]), [1, 1, 1], 156.25, [3, 3, 3])
    g64 = api.compute_grid(ctx, fw, [3, 3, 3], api.ER_FP64)
    assert np.isfinite(ref[0, 0, 0]) and ref[0, 0, 0] > 1e150
    assert abs(g64[0, 0, 0] - ref[0, 0, 0]) / ref[0, 0, 0] < 1e-9
    g32 = api.compute_grid(ctx, fw, [3, 3, 3], api.ER_FP32)
    assert g32[0, 0, 0] == np.float32(3.4028234663852886e38)
    fw.close()
