"""Batched guest-framework energy calls (er_grid_energies): the wrap of
src/gcmc.cc:74-77 and the interpolation of GuestFrameworkEnergy
(src/gcmc.cc:145-177) at caller-given positions, on FP32 and FP64 grids.

Both sides use separate multiplies and adds and IEEE division, so the
energies must equal the oracle's bit for bit, blocked (1e13 K) nodes
included; positions cover every wrap case (negative, above 1, exact
integers, grid nodes and the f = 1 face).
"""
import numpy as np
import pytest

from emptyroom_amd import api, synth
from oracle import oracle

pytestmark = pytest.mark.gpu


def positions(rng, n, shape):
    f = rng.uniform(-2.0, 3.0, (n, 3))
    nodes = np.stack([rng.integers(0, s, 64) / (s - 1.0) for s in shape], 1)
    special = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [-1.0, 2.0, -3.0], [0.5, 0.5, 0.5],
                        [1.0 - 2.0 ** -52, 2.0 ** -60, -2.0 ** -60]])
    return np.concatenate([special, nodes, f])


def test_fp32_grid_energies_bit_exact(ctx):
    print("grid energies FP32", flush=True)
    rng = np.random.default_rng(5)
    g = rng.normal(-600.0, 300.0, (37, 29, 41)).astype(np.float32)
    g[rng.random(g.shape) < 0.03] = 1e13
    G = api.Grid.upload(ctx, g)
    f = positions(rng, 20000, g.shape)
    got = api.grid_energies(ctx, G, f)
    G.close()
    ref = oracle.trilinear_points(g.astype(np.float64), f)
    assert np.array_equal(got, ref)
    G = api.Grid.upload(ctx, g)
    assert len(api.grid_energies(ctx, G, np.zeros((0, 3)))) == 0
    G.close()


def test_fp64_grid_energies_bit_exact(ctx):
    print("grid energies FP64", flush=True)
    ctx.tuning = api.default_tuning()
    fr = synth.small_framework(61, n_atoms=40)
    a = fr.particles()
    fw = api.Framework(ctx, a[:, :3], a[:, 3], a[:, 4], fr.t, fr.rep, fr.rc2)
    n = [19, 23, 17]
    G = api.Grid.compute(ctx, fw, n, api.ER_FP64)
    g = G.download()
    rng = np.random.default_rng(6)
    f = positions(rng, 5000, g.shape)
    got = api.grid_energies(ctx, G, f)
    G.close()
    fw.close()
    assert np.array_equal(got, oracle.trilinear_points(g, f))
