"""GPU parity of the two ways a Widom bin keeps its insertions' energies.

A bin whose insertions fit one chunk of the tiled kernel (at most 512: two
per lane) keeps their points and partial energies in registers across its
LDS fill passes; a larger bin stores and reloads them after every pass.
The insertions-per-bin knob (a mean; the bin count is n / knob) puts every
bin of these runs on one side: 64 per bin gives bins far below 512 (one
point per lane, some waves idle), 400 a mix of one and two points per lane,
4000 bins of several chunks.  Each run is checked against the oracle at the
same insertion points, with and without the saturation stop.

Tolerances as for every Widom test: energies within 1e-3 K where E < 1e4 K,
K_H within 1e-4 relative, <E> within 1e-3 K (relative beyond 1 K).
"""
import numpy as np
import pytest

from emptyroom_amd import api, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

CELL = (14.0, 15.0, 13.0, 80.0, 100.0, 95.0)  # ~40% of insertions below 1e4 K


@pytest.fixture(scope="module")
def wctx():
    c = api.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("per_bin", [64, 400, 4000])
def test_widom_bins_match_oracle(wctx, per_bin):
    print(f"widom bins per_bin={per_bin}", flush=True)
    t = wctx.tuning
    default = t.widom_per_bin
    t.widom_per_bin = per_bin
    wctx.tuning = t
    try:
        f = synth.small_framework(23, n_atoms=30, cell=CELL)
        a = f.particles()
        fw = api.Framework(wctx, a[:, :3], a[:, 3], a[:, 4], f.t, f.rep, f.rc2)
        nins, seed, T = 20000, 1234, 300.0
        sb, seb, n, e = api.widom(wctx, fw, nins, T, seed, energies=True)
        rb, reb, re = oracle.widom(f.t, a, f.rep, f.rc2, T, seed, 0, nins, energies=True)
        assert n == nins
        mask = re < 1e4
        assert mask.sum() > 100
        assert np.max(np.abs(e[mask] - re[mask])) <= 1e-3
        assert abs(sb - rb) / rb <= 1e-4
        assert abs(seb / sb - reb / rb) <= 1e-3 * max(1.0, abs(reb / rb))
        # the saturation stop (energies kept internal) leaves the sums as they are
        sb2, seb2, n2 = api.widom(wctx, fw, nins, T, seed)
        assert n2 == nins
        assert abs(sb2 - rb) / rb <= 1e-4
        assert abs(seb2 / sb2 - reb / rb) <= 1e-3 * max(1.0, abs(reb / rb))
        fw.close()
    finally:
        t.widom_per_bin = default
        wctx.tuning = t
