"""Grid points that sit exactly on a framework atom.

The reference's IRMOF-1 cell (25.832 A cubic) at GridResolution 0.15 A has
N = 173 points per axis, so (N - 1) / 4 = 43 and the grid point
(43, 43, 43) / 172 = (1/4, 1/4, 1/4) is the O atom at (1/4, 1/4, 1/4)
(src/writegrid.cu:122-127 forms both coordinates the same way).  There
r^2 = 0 exactly and the reference clamps it to min_r^2 = 1e-24
(src/computegridsheet.h:19,77), so the energy is about 4 eps sigma^12 1e144.

CPU: the oracle gives that clamped value at the coincident point (closed
form: the clamped pair plus the oracle's energy of the other atoms).
GPU: the FP64 grid equals the oracle at every coincident point and at 2000
random points to 1e-9 relative, and the FP32 grid holds FLT_MAX there.
"""
import os
import re
import shutil

import numpy as np
import pytest

from emptyroom_amd import api
from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "irmof1")


def setup_run(tmp_path):
    d = tmp_path / "run"
    shutil.copytree(GOLDEN, d, copy_function=shutil.copyfile)
    for root, dirs, files in os.walk(d):
        os.chmod(root, 0o755)
        for name in files:
            os.chmod(os.path.join(root, name), 0o644)
    text = (d / "simulation.input").read_text()
    text = re.sub(r"^GridResolution .*$", "GridResolution 0.15", text, flags=re.M)
    (d / "simulation.input").write_text(text)
    hr = api.HostRun(str(d), "IRMOF-1_clean", "Kr", 0)
    info = hr.info()
    atoms = hr.particles()
    hr.close()
    t = np.array(info.t_matrix[:]).reshape(3, 3)
    rep = list(info.rep[:])
    n = list(info.grid_n[:])
    return atoms, t, rep, info.r_cutoff_squared, n


def coincident(atoms, n):
    """Grid indices of the points that equal an atom's fractional coordinates."""
    out = []
    for a in atoms:
        idx = []
        for d in range(3):
            i = int(round(a[d] * (n[d] - 1)))
            if not (0 <= i < n[d] and 1.0 * i / (n[d] - 1) == a[d]):
                break
            idx.append(i)
        else:
            out.append(idx)
    return np.array(out, dtype=np.int64).reshape(-1, 3)


def test_oracle_clamps_a_point_on_an_atom(tmp_path):
    atoms, t, rep, rc2, n = setup_run(tmp_path)
    assert n == [173, 173, 173]
    idx = coincident(atoms, n)
    assert len(idx) >= 1
    on = [k for k, a in enumerate(atoms)
          if np.array_equal(a[:3], idx[0] / (np.array(n) - 1.0))]
    k = on[0]
    f = idx[0] / (np.array(n) - 1.0)
    e = oracle.lj_points(t, atoms, rep, rc2, 1e-24, f[None, :])[0]
    others = np.delete(atoms, k, axis=0)
    e_rest = oracle.lj_points(t, others, rep, rc2, 1e-24, f[None, :])[0]
    # the atom's own images: the one at the point (clamped r^2 = 1e-24) plus
    # the others, which sit a whole lattice vector away
    eps, sig2 = atoms[k, 3], atoms[k, 4]
    s6 = (sig2 / 1e-24) ** 3
    clamped = 4.0 * eps * s6 * (s6 - 1.0)
    img = oracle.lj_points(t, atoms[k:k + 1], [0, 0, 0], rc2, 1e-24, f[None, :])[0]
    assert img == clamped
    assert np.isfinite(e) and e > 1e140
    assert abs(e - (clamped + (oracle.lj_points(t, atoms[k:k + 1], rep, rc2, 1e-24,
                                                f[None, :])[0] - clamped) + e_rest)) \
        <= 1e-12 * e


@pytest.mark.gpu
def test_fp64_grid_on_atoms_matches_oracle(ctx, tmp_path):
    atoms, t, rep, rc2, n = setup_run(tmp_path)
    print(f"IRMOF-1 Kr grid {n} with points on atoms", flush=True)
    ctx.tuning = api.default_tuning()
    fw = api.Framework(ctx, atoms[:, :3], atoms[:, 3], atoms[:, 4], t, rep, rc2)
    g64 = api.compute_grid(ctx, fw, n, api.ER_FP64)
    g32 = api.compute_grid(ctx, fw, n, api.ER_FP32)
    fw.close()
    idx = coincident(atoms, n)
    rng = np.random.default_rng(9)
    rnd = np.stack([rng.integers(0, n[d], 2000) for d in range(3)], 1)
    allidx = np.concatenate([idx, rnd])
    ref = oracle.lj_points(t, atoms, rep, rc2, 1e-24, allidx / (np.array(n) - 1.0))
    got = g64[allidx[:, 0], allidx[:, 1], allidx[:, 2]]
    rel = np.abs(got - ref) / np.maximum(1.0, np.abs(ref))
    assert rel.max() < 1e-9, rel.max()
    assert np.all(ref[:len(idx)] > 1e140)
    assert np.all(g32[idx[:, 0], idx[:, 1], idx[:, 2]] == np.finfo(np.float32).max)
