"""Host side of the drop-ins vs the reference's own host code (CPU).

oracle/_ref/ref_host is built from the reference's Framework.cc,
Forcefield.cc, readsettings.h, load_fast_particle_f_array.h and
write_settings_to_outputfile.h (oracle/Makefile).  The product's loaders
(emptyroom_amd/csrc/er_host.cpp, through the C ABI) must load the same values
bit for bit, write the same .out headers, and fail with the same messages.
The IRMOF-1 known answers are also checked as constants, so these tests keep
their meaning where the reference tree is absent.
"""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from emptyroom_amd import api
from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "irmof1")
needs_ref = pytest.mark.skipif(not oracle.have_ref(), reason="oracle/_ref not built")


def workdir(tmp_path):
    d = tmp_path / "run"
    shutil.copytree(GOLDEN, d, copy_function=shutil.copyfile)
    for root, dirs, files in os.walk(d):  # the checkout may be read-only
        os.chmod(root, 0o755)
        for name in files:
            os.chmod(os.path.join(root, name), 0o644)
    (d / "output_files").mkdir()
    (d / "data" / "grids").mkdir()
    return d


def ref_dump(d, mode, structure, adsorbate):
    p = subprocess.run([oracle.REF_HOST, mode, structure, adsorbate], cwd=d,
                       capture_output=True, text=True)
    vals, atoms = {}, []
    for line in p.stdout.splitlines():
        w = line.split()
        if not w:
            continue
        if w[0] == "atom":
            atoms.append((w[2], [float(x) for x in w[3:9]]))
        elif w[0] in ("forcefield", "gridoutputformat"):
            vals[w[0]] = w[1]
        else:
            try:
                vals[w[0]] = [float(x) for x in w[1:]]
            except ValueError:
                pass
    return p.returncode, p.stdout, vals, atoms


def test_irmof1_known_answers(tmp_path):
    d = workdir(tmp_path)
    r = api.HostRun(str(d), "IRMOF-1_clean", "Xe", 1)
    i = r.info()
    assert i.n_atoms == 424
    assert f"{i.density:f}" == "593.306519"
    assert i.a == i.b == i.c == 25.832
    assert all(abs(math.degrees(x) - 90.0) < 1e-12 for x in (i.alpha, i.beta, i.gamma))
    t = np.array(i.t_matrix).reshape(3, 3)
    assert np.allclose(t, np.diag([25.832] * 3), atol=1e-12)
    labels = r.labels()
    assert {k: labels.count(k) for k in set(labels)} == {"C": 192, "H": 96, "O": 104, "Zn": 32}
    assert (i.numinsertions, i.insertions_per_thread, i.total_insertions) == \
        (17237492, 1053, 17252352)
    assert list(i.rep) == [1, 1, 1] and i.r_cutoff_squared == 156.25 and i.temperature == 298.0
    r.close()


def write_synthetic_inputs(d):
    """A triclinic structure with coordinates outside [0,1), several labels, a
    forcefield with a blank line and a guest defined last."""
    (d / "data" / "structures" / "tric.cssr").write_text(
        "  11.5 12.25 13.75\n  78.5 101.25 96.0 SPGR = 1 P 1\n6 0\n0 tric : tric\n"
        " 1 Zn 0.10 0.20 0.30 0 0\n 2 O -0.25 1.40 0.50\n 3 C 2.75 -1.10 0.00\n"
        " 4 H 0.999 0.001 -0.5\n 5 O 0.5 0.5 0.5\n 6 C -0.0 0.75 1.0\n")
    (d / "data" / "forcefields" / "toy.def").write_text(
        "[Atom] [eps] [sig]\nZn 62.4 2.46\nO 30.2 3.12\n\nC 52.8 3.43\nH 22.1 2.57\n"
        "Xe 221.0 4.1\n")
    (d / "data" / "uc_replications" / "tric_once.uc").write_text("3 3 2")
    (d / "simulation.input").write_text(
        "Forcefield toy other.def\nGridResolution 0.7 A\nCutoffRadius 11.0\n"
        "Temperature 310 K\nNumberOfInsertionsPerA3 20\nGridOutputFormat cube\n")


@needs_ref
@pytest.mark.parametrize("structure,adsorbate,synthetic",
                         [("IRMOF-1_clean", "Xe", False), ("IRMOF-1_clean", "Kr", False),
                          ("tric", "Xe", True)])
@pytest.mark.parametrize("mode", ["grid", "henry"])
def test_loaders_bit_identical_to_reference(tmp_path, structure, adsorbate, synthetic, mode):
    d = workdir(tmp_path)
    if synthetic:
        write_synthetic_inputs(d)
    rc, out, ref, ref_atoms = ref_dump(d, mode, structure, adsorbate)
    assert rc == 0, out
    r = api.HostRun(str(d), structure, adsorbate, 0 if mode == "grid" else 1)
    i = r.info()
    assert [i.a, i.b, i.c, i.alpha, i.beta, i.gamma] == ref["cell"]
    assert [i.density] == ref["density"] and [i.volume] == ref["volume"]
    assert list(i.t_matrix) == ref["t_matrix"] and list(i.inv_t_matrix) == ref["inv_t_matrix"]
    assert [float(x) for x in i.rep] == ref["rep"]
    assert [i.r_cutoff_squared] == ref["r_cutoff_squared"]
    assert [i.epsilon_guest, i.sigma_guest] == ref["guest"]
    assert [i.adsorbate_mw] == ref["adsorbate_mw"]
    if mode == "grid":
        assert [float(x) for x in i.grid_n] == ref["grid_n"]
    else:
        assert [float(i.numinsertions)] == ref["numinsertions"]
        assert [float(i.insertions_per_thread)] == ref["per_thread"]
    P = r.particles()
    labels = r.labels()
    assert len(P) == len(ref_atoms)
    for k, (lab, v) in enumerate(ref_atoms):
        assert labels[k] == lab
        assert list(P[k, :3]) == v[:3]           # fmod-wrapped fractional coordinates
        assert P[k, 3] == v[4] and P[k, 4] == v[5]  # mixed eps, sigma^2
    ours = d / "ours.out"
    r.write_settings(str(ours))
    refout = d / "output_files" / f"{structure}_{adsorbate}_{mode}.out"
    a = ours.read_text().splitlines()
    b = refout.read_text().splitlines()
    assert a[0].startswith("date: ") and b[0].startswith("date: ")
    assert a[1:] == b[1:]
    r.close()


ERROR_CASES = [
    # (description, mutation, mode, structure, adsorbate)
    ("guest missing from masses.def", None, "grid", "IRMOF-1_clean", "CH4"),
    ("no masses.def", "rm data/masses.def", "henry", "IRMOF-1_clean", "Xe"),
    ("no simulation.input", "rm simulation.input", "grid", "IRMOF-1_clean", "Xe"),
    ("no GridResolution", "sed GridResolution", "grid", "IRMOF-1_clean", "Xe"),
    ("no NumberOfInsertionsPerA3", "sed NumberOfInsertionsPerA3", "henry", "IRMOF-1_clean",
     "Xe"),
    ("no Temperature (henry)", "sed Temperature", "henry", "IRMOF-1_clean", "Xe"),
    ("no Forcefield", "sed Forcefield", "grid", "IRMOF-1_clean", "Xe"),
    ("no .uc file", "rm data/uc_replications/IRMOF-1_clean_once.uc", "grid", "IRMOF-1_clean",
     "Xe"),
    ("bad .uc file", "write data/uc_replications/IRMOF-1_clean_once.uc=1 x", "henry",
     "IRMOF-1_clean", "Xe"),
    ("no forcefield file", "rm data/forcefields/UFF.def", "grid", "IRMOF-1_clean", "Xe"),
    ("no cssr", None, "grid", "NOPE", "Xe"),
    ("atom without mass", "sed-masses Zn", "henry", "IRMOF-1_clean", "Xe"),
    ("guest not in forcefield", "append-masses Qq 12.0", "grid", "IRMOF-1_clean", "Qq"),
    ("atom not in forcefield", "sed-ff Zn", "grid", "IRMOF-1_clean", "Xe"),
]


def mutate(d, m):
    if m is None:
        return
    if m.startswith("rm "):
        os.remove(d / m[3:])
    elif m.startswith("write "):
        path, text = m[6:].split("=", 1)
        (d / path).write_text(text)
    elif m.startswith("sed-masses "):
        p = d / "data" / "masses.def"
        p.write_text("\n".join(l for l in p.read_text().splitlines()
                               if not l.startswith(m[11:] + " ")) + "\n")
    elif m.startswith("append-masses "):
        p = d / "data" / "masses.def"
        p.write_text(p.read_text().rstrip("\n") + "\n" + m[14:] + "\n")
    elif m.startswith("sed-ff "):
        p = d / "data" / "forcefields" / "UFF.def"
        p.write_text("\n".join(l for l in p.read_text().splitlines()
                               if not l.startswith(m[7:] + " ")) + "\n")
    elif m.startswith("sed "):
        p = d / "simulation.input"
        p.write_text("\n".join(l for l in p.read_text().splitlines()
                               if not l.startswith(m[4:])) + "\n")


@needs_ref
@pytest.mark.parametrize("desc,m,mode,structure,adsorbate", ERROR_CASES,
                         ids=[c[0] for c in ERROR_CASES])
def test_error_messages_match_reference(tmp_path, desc, m, mode, structure, adsorbate):
    d = workdir(tmp_path)
    mutate(d, m)
    rc, out, _, _ = ref_dump(d, mode, structure, adsorbate)
    assert rc != 0
    with pytest.raises(api.EmptyRoomError) as e:
        api.HostRun(str(d), structure, adsorbate, 0 if mode == "grid" else 1)
    assert e.value.message == out
    # the drop-in binary prints the same text and fails the same way
    exe = os.path.join(api.BIN_DIR, "writegrid" if mode == "grid" else "henry")
    p = subprocess.run([exe, structure, adsorbate], cwd=d, capture_output=True, text=True)
    assert p.returncode != 0 and p.stdout == out


def test_usage_message(tmp_path):
    for exe, name in (("writegrid", "writegrid"), ("henry", "henry")):
        p = subprocess.run([os.path.join(api.BIN_DIR, exe)], capture_output=True, text=True)
        assert p.returncode != 0
        assert p.stdout == f"Run as:\n./{name} structure_name AdsorbateID\n"


def test_grid_file_formats(tmp_path):
    """txt and cube layouts of src/writegrid.cu:147-167,244-269."""
    rng = np.random.default_rng(0)
    g = rng.normal(0, 1e3, size=(3, 4, 13))
    g[0, 0, 0] = 1.25e152
    g[1, 2, 3] = -0.0
    t = np.array([[10.0, 1.5, -2.0], [0.0, 11.0, 0.7], [0.0, 0.0, 12.0]])
    txt = tmp_path / "g.txt"
    api.write_grid_file(str(txt), "txt", g, t, threads=3)
    lines = [f"{3} {4} {13}  = (N_x,N_y,N_z) grid points (grid is in fractional coords). "
             "Endpoints included."]
    for i in range(3):
        for j in range(4):
            lines.append("".join("% 13.6E " % v for v in g[i, j]))
    assert txt.read_text() == "\n".join(lines) + "\n"
    cube = tmp_path / "g.cube"
    g = np.clip(g, -1e30, 1e30)  # the cube check runs on an FP32 grid
    api.write_grid_file(str(cube), "cube", g.astype(np.float32), t, threads=2)
    n = g.shape
    exp = "\nThis is a grid file.\n"
    exp += "%d % 13.6lf % 13.6lf % 13.6lf\n" % (0, 0.0, 0.0, 0.0)
    exp += "%d % 13.6lf % 13.6lf % 13.6lf\n" % (n[0], t[0, 0] / (n[0] - 1), 0.0, 0.0)
    exp += "%d % 13.6lf % 13.6lf % 13.6lf\n" % (n[1], t[0, 1] / (n[1] - 1), t[1, 1] / (n[1] - 1),
                                                0.0)
    exp += "%d % 13.6lf % 13.6lf % 13.6lf\n" % (n[2], t[0, 2] / (n[2] - 1), t[1, 2] / (n[2] - 1),
                                                t[2, 2] / (n[2] - 1))
    g32 = g.astype(np.float32).astype(np.float64)
    for i in range(n[0]):
        for j in range(n[1]):
            for k in range(n[2]):
                exp += "% 13.6E " % (g32[i, j, k] * 8.314 / 1000)
                if (k + 1) % 6 == 0:
                    exp += "\n"
            exp += "\n"
    assert cube.read_text() == exp
