"""The reference's direct method on the GPU (er_direct_grid / er_direct_widom).

It performs the reference's arithmetic (src/computegridsheet.h:54-87,
src/doHenry.h:69-99) in FP64, in the reference's order, with implicit FMA
contraction off -- the operations of the oracle's energy_from_images -- so
grids and per-insertion energies must equal the oracle's bit for bit.  The
Widom sums use the tile kernel's fixed-order reduction (a different order
from the oracle's sequential sum), so they agree to 1e-12 relative.
"""
import numpy as np
import pytest

from emptyroom_amd import api, synth
from oracle import oracle

pytestmark = pytest.mark.gpu


def fw_args(f):
    a = f.particles()
    return a, (a[:, :3], a[:, 3], a[:, 4], f.t, f.rep, f.rc2)


@pytest.mark.parametrize("seed", [11, 12])
def test_direct_grid_and_widom_bit_exact(ctx, seed):
    print(f"direct method, triclinic seed {seed}", flush=True)
    f = synth.small_framework(seed, n_atoms=40)
    a, args = fw_args(f)
    n = [13, 17, 11]
    g = api.direct_grid(ctx, *args, n)
    ref = oracle.lj_grid(f.t, a, f.rep, f.rc2, n)
    assert np.array_equal(g, ref)
    nins, T, s = 30000, 298.0, 5
    sb, seb, cnt, e = api.direct_widom(ctx, *args, nins, T, s, energies=True)
    rb, reb, re = oracle.widom(f.t, a, f.rep, f.rc2, T, s, 0, nins, energies=True)
    assert cnt == nins
    assert np.array_equal(e, re)
    assert abs(sb - rb) <= 1e-12 * rb
    assert abs(seb - reb) <= 1e-12 * np.abs(re * np.exp(-re / T)).sum()


def test_direct_config1_and_tile_kernel(ctx):
    """Config 1 (cubic 25 A, 400 atoms, 50^3 points, 1e5 insertions): the
    direct method equals the oracle; the tile kernel's FP64 grid agrees with
    it to 1e-9 relative and its Widom sums to 1e-4."""
    print("direct method, config 1", flush=True)
    f = synth.config1()
    a, args = fw_args(f)
    n = f.grid_n()
    g = api.direct_grid(ctx, *args, n)
    assert np.array_equal(g, oracle.lj_grid(f.t, a, f.rep, f.rc2, n))
    ctx.tuning = api.default_tuning()
    fw = api.Framework(ctx, *args)
    g64 = api.compute_grid(ctx, fw, n, api.ER_FP64)
    rel = np.abs(g64 - g) / np.maximum(1.0, np.abs(g))
    assert rel.max() < 1e-9, rel.max()
    nins, T, s = 100000, 298.0, 9
    sb, seb, _ = api.direct_widom(ctx, *args, nins, T, s)
    pb, peb, _ = api.widom(ctx, fw, nins, T, s)
    fw.close()
    assert abs(pb - sb) / sb <= 1e-4
    rb, _ = oracle.widom(f.t, a, f.rep, f.rc2, T, s, 0, nins)
    assert abs(sb - rb) <= 1e-12 * rb


def test_direct_edge_cases(ctx):
    print("direct method, edge cases", flush=True)
    t = np.diag([10.0, 11.0, 12.0])
    empty = np.zeros((0, 3))
    z = np.zeros(0)
    g = api.direct_grid(ctx, empty, z, z, t, [1, 1, 1], 12.5 ** 2, [2, 3, 4])
    assert g.shape == (2, 3, 4) and not g.any()
    sb, seb, cnt = api.direct_widom(ctx, empty, z, z, t, [1, 1, 1], 12.5 ** 2, 1000, 300.0, 1)
    assert (sb, seb, cnt) == (1000.0, 0.0, 1000)
    # one atom on a grid node: the clamp of src/computegridsheet.h:77
    one = np.array([[0.5, 0.5, 0.5]])
    eps, sig2 = np.array([100.0]), np.array([9.0])
    g = api.direct_grid(ctx, one, eps, sig2, t, [0, 0, 0], 12.5 ** 2, [3, 3, 3])
    ref = oracle.lj_grid(t, np.column_stack([one, eps, sig2]), [0, 0, 0], 12.5 ** 2, [3, 3, 3])
    assert np.array_equal(g, ref) and g[1, 1, 1] > 1e100
    with pytest.raises(api.EmptyRoomError):
        api.direct_grid(ctx, one, eps, sig2, t, [0, 0, 0], 12.5 ** 2, [2048, 2048, 1024])
    with pytest.raises(api.EmptyRoomError):
        api.direct_widom(ctx, one, eps, sig2, t, [0, 0, 0], 12.5 ** 2, 0, 300.0, 1)
