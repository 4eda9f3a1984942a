"""GPU parity for frameworks with many pair types.

Each pair type is one segment of the LDS list, a pass holds at most 64
segments, and region sub-lists keep per-type offsets for at most 64 types
(more types stream the whole near list).  40 types exercise many segments
per pass with region sub-lists; 70 types exceed both limits (passes end at
the 64th segment, regions fall back to the near list).  A dense cluster
makes several passes per tile.  Tolerances as in the other parity tests.
"""
import numpy as np
import pytest

from emptyroom_amd import api, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

def check_fp32(g32, ref):
    mask = np.abs(ref) < 1e4
    if mask.any():
        err = float(np.max(np.abs(g32.astype(np.float64)[mask] - ref[mask])))
        assert err <= 1e-3, f"max abs error {err} K over |E| < 1e4 K"
    hi = ~mask & (np.abs(ref) < 3.0e38)
    if hi.any():
        rel = np.abs(g32.astype(np.float64)[hi] - ref[hi]) / np.abs(ref[hi])
        assert rel.max() < 1e-6, rel.max()


def many_type_atoms(ntypes, seed):
    """One atom per pair type (CH4 guest, Lorentz-Berthelot mixing), about
    0.02 atoms / A^3 so that many insertions are accessible."""
    rng = np.random.default_rng(seed)
    edge = 12.0 * (ntypes / 40.0) ** (1.0 / 3.0)
    cell = (edge, edge + 0.5, edge + 1.0, 85.0, 95.0, 100.0)
    t, _ = synth.cell_matrix(*cell)
    eps_i = rng.uniform(20.0, 250.0, ntypes)
    sig_i = rng.uniform(2.5, 4.0, ntypes)
    eps = np.sqrt(148.0 * eps_i)
    sig2 = ((3.73 + sig_i) / 2.0) ** 2
    frac = rng.random((ntypes, 3))
    return t, np.column_stack([frac, eps, sig2]), synth.replication(t, 12.5)


@pytest.mark.parametrize("ntypes", [40, 70])
def test_many_types_grid_and_widom(ctx, ntypes):
    print(f"many types: {ntypes}", flush=True)
    ctx.tuning = api.default_tuning()
    t, a, rep = many_type_atoms(ntypes, 500 + ntypes)
    assert len(np.unique(a[:, 3:5], axis=0)) == ntypes
    rc2 = 12.5 ** 2
    fw = api.Framework(ctx, a[:, :3], a[:, 3], a[:, 4], t, rep, rc2)
    n = [41, 43, 45]  # ~0.3 A: regions of several tiles form
    ref = oracle.lj_grid(t, a, rep, rc2, n)
    g64 = api.compute_grid(ctx, fw, n, api.ER_FP64)
    rel = np.abs(g64 - ref) / np.maximum(1.0, np.abs(ref))
    assert rel.max() < 1e-9, rel.max()
    check_fp32(api.compute_grid(ctx, fw, n, api.ER_FP32), ref)
    assert api.count_incut_grid(ctx, fw, n) == oracle.count_incut_grid(t, a, rep, rc2, n)
    nins, T, seed = 20000, 298.0, 8
    sb, seb, cnt, e = api.widom(ctx, fw, nins, T, seed, energies=True)
    rb, reb, re = oracle.widom(t, a, rep, rc2, T, seed, 0, nins, energies=True)
    m = re < 1e4
    assert m.sum() > 1000, m.sum()
    assert np.max(np.abs(e[m] - re[m])) <= 1e-3
    assert abs(sb - rb) / rb <= 1e-4
    fw.close()


def test_dense_cluster_many_passes(ctx):
    """600 atoms of 12 types in a 6 A ball: every tile's list holds far more
    records than one LDS pass takes."""
    print("dense cluster, many passes", flush=True)
    ctx.tuning = api.default_tuning()
    rng = np.random.default_rng(4)
    L = 30.0
    t = np.diag([L, L, L])
    d = rng.normal(size=(600, 3))
    d *= (6.0 / L) * (rng.random(600) ** (1 / 3) / np.linalg.norm(d, axis=1))[:, None]
    frac = (0.5 + d) % 1.0
    eps = rng.choice(np.linspace(30.0, 200.0, 12), 600)
    sig2 = rng.choice(np.linspace(3.0, 3.6, 12), 600) ** 2
    a = np.column_stack([frac, eps, sig2])
    rc2 = 12.5 ** 2
    fw = api.Framework(ctx, frac, eps, sig2, t, [1, 1, 1], rc2)
    n = [31, 31, 31]
    ref = oracle.lj_grid(t, a, [1, 1, 1], rc2, n)
    g64 = api.compute_grid(ctx, fw, n, api.ER_FP64)
    rel = np.abs(g64 - ref) / np.maximum(1.0, np.abs(ref))
    assert rel.max() < 1e-9, rel.max()
    check_fp32(api.compute_grid(ctx, fw, n, api.ER_FP32), ref)
    assert api.count_incut_grid(ctx, fw, n) == oracle.count_incut_grid(t, a, [1, 1, 1], rc2, n)
    fw.close()
