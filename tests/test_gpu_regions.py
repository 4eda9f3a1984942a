"""GPU parity of region sub-lists in the tiled grid kernel.

A grid work item is a region of tiles: the workgroup first gathers every
image atom within r_c + R of the region centre (R bounds every tile's own
sphere), then fills each tile's LDS list from that sub-list.  Regions only
form on fine grids (tile edge well below the region length), so these tests
use 0.2-0.3 A spacings, odd point counts (partial regions at the far faces),
several region lengths, and a clustered framework whose sub-list overflows
the per-workgroup scratch and falls back to the full near list.

Tolerances: FP64 grid 1e-9 relative to the oracle; FP32 grid 1e-3 K absolute
where |E| < 1e4 K and 1e-6 relative beyond (the clustered framework reaches
E ~ -3e4 K, where FP32 storage alone rounds by up to 2e-3 K); counts exact.
"""
import numpy as np
import pytest

from emptyroom_amd import api, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

CELL = (10.0, 11.0, 12.0, 85.0, 95.0, 100.0)


@pytest.fixture(scope="module")
def rctx():
    c = api.Context(0)
    yield c
    c.close()


def set_region(c, length):
    t = c.tuning
    t.region_len = length
    c.tuning = t


def upload(c, frac, eps, sig2, t, rep, rc2):
    return api.Framework(c, frac, eps, sig2, t, rep, rc2)


def check_fp32(g32, ref):
    # 1e-3 K where |E| < 1e4 K; beyond that an FP32 value cannot hold 1e-3 K
    # (half an ulp is 9.8e-4 K at 16384 K), so the bound is relative there.
    mask = np.abs(ref) < 1e4
    if mask.any():
        err = float(np.max(np.abs(g32.astype(np.float64)[mask] - ref[mask])))
        assert err <= 1e-3, f"max abs error {err} K over |E| < 1e4 K"
    hi = ~mask & (np.abs(ref) < 3.0e38)
    if hi.any():
        rel = np.abs(g32.astype(np.float64)[hi] - ref[hi]) / np.abs(ref[hi])
        assert rel.max() < 1e-6, rel.max()


@pytest.mark.parametrize("region_len", [2.5, 4.0, 100.0])
def test_fine_grid_regions_match_oracle(rctx, region_len):
    print(f"regions fine grid region_len={region_len}", flush=True)
    set_region(rctx, region_len)
    f = synth.small_framework(41, n_atoms=70, cell=CELL)
    a = f.particles()
    n = [41, 47, 53]  # about 0.25 A; tile counts not divisible by the region extents
    fw = upload(rctx, a[:, :3], a[:, 3], a[:, 4], f.t, f.rep, f.rc2)
    ref = oracle.lj_grid(f.t, a, f.rep, f.rc2, n)
    g64 = api.compute_grid(rctx, fw, n, api.ER_FP64)
    rel = np.abs(g64 - ref) / np.maximum(1.0, np.abs(ref))
    assert rel.max() < 1e-9, rel.max()
    check_fp32(api.compute_grid(rctx, fw, n, api.ER_FP32), ref)
    assert api.count_incut_grid(rctx, fw, n) == oracle.count_incut_grid(f.t, a, f.rep, f.rc2, n)
    fw.close()
    set_region(rctx, 4.0)


def test_regions_off_and_on_agree(rctx):
    f = synth.small_framework(42, n_atoms=80, cell=(9.0, 14.0, 8.5, 70.0, 110.0, 65.0))
    a = f.particles()
    n = [37, 57, 35]
    out = {}
    for length in (0.0, 3.0):
        set_region(rctx, length)
        fw = upload(rctx, a[:, :3], a[:, 3], a[:, 4], f.t, f.rep, f.rc2)
        out[length] = (api.compute_grid(rctx, fw, n, api.ER_FP64),
                       api.compute_grid(rctx, fw, n, api.ER_FP32))
        fw.close()
    set_region(rctx, 4.0)
    # a fill pass takes kept records up to the first that does not fit, so a
    # tile builds the same passes from its region's sub-list as from the near
    # list: bit-identical grids
    assert np.array_equal(out[0.0][0], out[3.0][0])
    assert np.array_equal(out[0.0][1], out[3.0][1])


def test_region_batch_and_server_bit_identical(rctx):
    set_region(rctx, 4.0)
    fs = [synth.small_framework(50 + s, n_atoms=40 + 15 * s, cell=CELL) for s in range(3)]
    fws = [upload(rctx, *(lambda a: (a[:, :3], a[:, 3], a[:, 4]))(f.particles()), f.t, f.rep,
                  f.rc2) for f in fs]
    tasks = [dict(fw=fw, grid_n=[33 + 4 * i, 45, 41], precision=api.ER_FP32,
                  n_insertions=3000, temperature=300.0, seed=5 + i) for i, fw in enumerate(fws)]
    plain = [api.compute_grid(rctx, t["fw"], t["grid_n"], api.ER_FP32) for t in tasks]
    b = api.Batch(rctx, tasks, use_graph=True)
    for _ in range(2):
        b.run()
        b.wait()
        for i in range(len(tasks)):
            assert np.array_equal(b.grid(i), plain[i])
    b.close()
    srv = api.Server(rctx, streams=2, items_per_launch=64)
    tickets = []
    for f, t in zip(fs, tasks):
        a = f.particles()
        code, tk = srv.submit(a[:, :3], a[:, 3], a[:, 4], f.t, f.rep, f.rc2, grid_n=t["grid_n"])
        assert code == 0
        tickets.append(tk)
    for i, tk in enumerate(tickets):
        assert srv.poll(tk, wait_ms=-1).state == api.RQ_DONE
        assert np.array_equal(srv.fetch_grid(tk, plain[i].shape), plain[i])
        srv.release(tk)
    srv.close()
    for fw in fws:
        fw.close()


def test_region_overflow_falls_back(rctx):
    """9000 atoms packed in a 5 A cube of a 60 A cell: the cell-average density
    lets regions form, but the sub-list of a region next to the cluster exceeds
    the 8192-record scratch, so that region streams the full near list."""
    print("regions overflow fallback", flush=True)
    set_region(rctx, 16.0)
    rng = np.random.default_rng(7)
    L = 60.0
    t = np.diag([L, L, L])
    frac = 0.45 + rng.random((9000, 3)) * (5.0 / L)
    eps = np.full(9000, 30.0)
    sig2 = np.full(9000, 9.0)
    a = np.column_stack([frac, eps, sig2])
    rc2 = 12.5 ** 2
    n = [61, 61, 61]
    fw = upload(rctx, frac, eps, sig2, t, [1, 1, 1], rc2)
    g64 = api.compute_grid(rctx, fw, n, api.ER_FP64)
    g32 = api.compute_grid(rctx, fw, n, api.ER_FP32)
    cnt = api.count_incut_grid(rctx, fw, n)
    fw.close()
    set_region(rctx, 0.0)
    fw = upload(rctx, frac, eps, sig2, t, [1, 1, 1], rc2)
    g64_full = api.compute_grid(rctx, fw, n, api.ER_FP64)
    assert api.count_incut_grid(rctx, fw, n) == cnt
    fw.close()
    set_region(rctx, 4.0)
    assert (np.abs(g64 - g64_full) / np.maximum(1.0, np.abs(g64_full))).max() < 1e-12
    # oracle at grid points around the cluster and far from it
    idx = np.stack(np.meshgrid(np.arange(22, 40, 3), np.arange(22, 40, 3), np.arange(0, 61, 4),
                               indexing="ij"), -1).reshape(-1, 3)
    fpts = idx / 60.0
    ref = oracle.lj_points(t, a, [1, 1, 1], rc2, 1e-24, fpts)
    got = g64[idx[:, 0], idx[:, 1], idx[:, 2]]
    assert (np.abs(got - ref) / np.maximum(1.0, np.abs(ref))).max() < 1e-9
    check_fp32(g32[idx[:, 0], idx[:, 1], idx[:, 2]], ref)


def test_widom_regions_batch_bit_identical_to_plain(rctx):
    """Widom bins of a batch are grouped into regions that share one sub-list
    (a batch of 32 jobs needs only 2048 / 32 items per job), while a plain
    run keeps one bin per item; a bin's candidate list, and so every energy
    and both sums, must be the same either way."""
    print("widom regions batch vs plain", flush=True)
    rctx.tuning = api.default_tuning()
    fs = [synth.small_framework(300 + s, n_atoms=40 + 3 * s,
                                cell=(11.0 + 0.2 * s, 12.0, 13.0, 85.0, 95.0, 100.0))
          for s in range(32)]
    fws = [upload(rctx, *(lambda a: (a[:, :3], a[:, 3], a[:, 4]))(f.particles()), f.t, f.rep,
                  f.rc2) for f in fs]
    tasks = [dict(fw=fw, grid_n=[0, 0, 0], precision=api.ER_FP32, n_insertions=300000,
                  temperature=298.0, seed=40 + i) for i, fw in enumerate(fws)]
    b = api.Batch(rctx, tasks, use_graph=False)
    b.run()
    b.wait()
    for i, t in enumerate(tasks):
        assert b.widom(i) == api.widom(rctx, t["fw"], t["n_insertions"], 298.0, t["seed"]), i
    b.close()
    # and the plain energies against the oracle on the first insertions
    f, fw = fs[5], fws[5]
    a = f.particles()
    _, _, _, e = api.widom(rctx, fw, 300000, 298.0, 45, energies=True)
    _, _, re = oracle.widom(f.t, a, f.rep, f.rc2, 298.0, 45, 0, 4000, energies=True)
    mask = re < 1e4
    assert np.max(np.abs(e[:4000][mask] - re[mask])) <= 1e-3
    for fw in fws:
        fw.close()
