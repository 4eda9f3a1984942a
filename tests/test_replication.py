"""Replication-factor tool (julia_codes/framework.jl:162-227) on the CPU:
the library's er_replication_factors / er_write_uc_file against the
reference's own .uc files for IRMOF-1_clean, the closed form through
perpendicular widths, and the oracle's restatement on random triclinic cells.
No GPU call is made."""
import ctypes
import math
import os
import shutil

import numpy as np
import pytest

from emptyroom_amd import api, synth
from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "irmof1")


def lib_rep(t, cutoff, twice):
    rep = (ctypes.c_int * 3)()
    t = np.ascontiguousarray(np.asarray(t, dtype=np.float64).reshape(9))
    code = api.lib().er_replication_factors(t.ctypes.data_as(ctypes.c_void_p), float(cutoff),
                                            int(twice), rep)
    assert code == 0, api.lib().er_last_error()
    return list(rep)


def closed_form(t, cutoff, twice):
    """r = 1 + ceil(log2(2 r_c / w)) for w < 2 r_c (w = perpendicular width):
    the centre of a cell doubled k times sits 2^k w / 2 from the face."""
    rc = 2.0 * cutoff if twice else cutoff
    w = synth.perp_widths(np.asarray(t).reshape(3, 3))
    return [1 if wi >= 2 * rc else 1 + math.ceil(math.log2(2 * rc / wi)) for wi in w]


@pytest.mark.parametrize("twice", [False, True])
def test_irmof1_uc_files_match_reference(tmp_path, twice):
    root = tmp_path / "run"
    shutil.copytree(GOLDEN, root, copy_function=shutil.copyfile)
    name = "IRMOF-1_clean_" + ("twice" if twice else "once") + ".uc"
    expected = (root / "data" / "uc_replications" / name).read_bytes()
    (root / "data" / "uc_replications" / name).unlink()
    rep = (ctypes.c_int * 3)()
    code = api.lib().er_write_uc_file(str(root).encode(), b"IRMOF-1_clean", 12.5, int(twice), rep)
    assert code == 0, api.lib().er_last_error()
    assert (root / "data" / "uc_replications" / name).read_bytes() == expected
    assert list(rep) == ([2, 2, 2] if twice else [1, 1, 1])


def test_oracle_pinned_by_closed_form_and_golden():
    t = np.diag([25.832] * 3)
    assert oracle.replication_factors(t, 12.5) == [1, 1, 1]
    assert oracle.replication_factors(t, 12.5, twice=True) == [2, 2, 2]
    rng = np.random.default_rng(3)
    for _ in range(200):
        cell = (*rng.uniform(5.0, 40.0, 3), *rng.uniform(65.0, 115.0, 3))
        t, v = synth.cell_matrix(*cell)
        if not v > 0:
            continue
        rc = float(rng.uniform(4.0, 20.0))
        for twice in (False, True):
            ref = oracle.replication_factors(t, rc, twice)
            # stay off the boundaries where the closed form's log2 ties
            w = synth.perp_widths(t)
            r2 = 2 * (2 * rc if twice else rc)
            if np.any(np.abs(np.log2(r2 / w) - np.round(np.log2(r2 / w))) < 1e-9):
                continue
            assert ref == closed_form(t, rc, twice)


def test_library_matches_oracle_on_triclinic_cells():
    rng = np.random.default_rng(11)
    for _ in range(300):
        cell = (*rng.uniform(3.0, 45.0, 3), *rng.uniform(60.0, 120.0, 3))
        t, v = synth.cell_matrix(*cell)
        if not v > 0:
            continue
        rc = float(rng.uniform(2.0, 25.0))
        for twice in (False, True):
            assert lib_rep(t, rc, twice) == oracle.replication_factors(t, rc, twice)


def test_degenerate_cell_is_an_error():
    rep = (ctypes.c_int * 3)()
    t = np.zeros(9)
    assert api.lib().er_replication_factors(t.ctypes.data_as(ctypes.c_void_p), 12.5, 0, rep) != 0


def test_ucrep_binary(tmp_path):
    exe = os.path.join(os.path.dirname(HERE), "emptyroom_amd", "bin", "ucrep")
    root = tmp_path / "run"
    shutil.copytree(GOLDEN, root, copy_function=shutil.copyfile)
    import subprocess
    out = subprocess.run([exe, "IRMOF-1_clean", "12.5", "twice"], cwd=root, capture_output=True,
                         text=True)
    assert out.returncode == 0, out.stdout
    assert (root / "data" / "uc_replications" / "IRMOF-1_clean_twice.uc").read_text() == "2 2 2"
    bad = subprocess.run([exe, "IRMOF-1_clean", "12.5", "thrice"], cwd=root, capture_output=True,
                         text=True)
    assert bad.returncode != 0 and "not valid for once_or_twice" in bad.stdout
