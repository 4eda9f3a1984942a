"""ctypes wrapper of liber_oracle.so -- the CPU FP64 restatement of the
reference arithmetic (see er_oracle.h).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline, never by the product.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liber_oracle.so")
REF_HOST = os.path.join(HERE, "_ref", "ref_host")

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "liber_oracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.orc_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.orc_widom_point.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, _D]
        L.orc_widom_points.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64,
                                       ctypes.c_int64, _D, ctypes.c_int]
        L.orc_cell_matrices.argtypes = [ctypes.c_double] * 6 + [_D, _D, _D]
        L.orc_lb_mix.argtypes = [ctypes.c_double] * 4 + [_D, _D]
        L.orc_image_atoms.argtypes = [_D, _D, ctypes.c_int, _I, _D]
        L.orc_lj_energy.argtypes = [_D, _D, ctypes.c_int, _I, ctypes.c_double, ctypes.c_double, _D]
        L.orc_lj_energy.restype = ctypes.c_double
        L.orc_lj_grid.argtypes = [_D, _D, ctypes.c_int, _I, ctypes.c_double] + \
            [ctypes.c_int] * 6 + [_D, ctypes.c_int]
        L.orc_lj_points.argtypes = [_D, _D, ctypes.c_int, _I, ctypes.c_double, ctypes.c_double,
                                    _D, ctypes.c_int64, _D, ctypes.c_int]
        L.orc_accessibility_grid.argtypes = [_D, _D, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, _D, ctypes.c_int]
        L.orc_count_incut_grid.argtypes = [_D, _D, ctypes.c_int, _I, ctypes.c_double] + \
            [ctypes.c_int] * 6 + [ctypes.c_int]
        L.orc_count_incut_grid.restype = ctypes.c_int64
        L.orc_widom.argtypes = [_D, _D, ctypes.c_int, _I, ctypes.c_double, ctypes.c_double,
                                ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, _D, _D, _D,
                                ctypes.c_int]
        L.orc_henry_counts.argtypes = [ctypes.c_int, ctypes.c_double, _I, _I,
                                       ctypes.POINTER(ctypes.c_int64)]
        L.orc_henry_result.argtypes = [ctypes.c_double] * 5 + [_D] * 4
        L.orc_find_and_block_pockets.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_double, ctypes.POINTER(ctypes.c_int32)]
        L.orc_find_and_block_pockets.restype = ctypes.c_int
        L.orc_wrap.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_wrap.restype = ctypes.c_double
        L.orc_trilinear.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int] + \
            [ctypes.c_double] * 3
        L.orc_trilinear.restype = ctypes.c_double
        L.orc_widom_grid.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, _D, _D,
                                     _D, ctypes.c_int]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_D)


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def _rep(rep):
    return (ctypes.c_int * 3)(*[int(x) for x in rep])


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def widom_point(seed: int, stream: int, i: int) -> np.ndarray:
    f = np.empty(3)
    lib().orc_widom_point(seed, stream, i, _d(f))
    return f


def widom_points(seed: int, stream: int, i0: int, n: int, nthreads=0) -> np.ndarray:
    """Insertion points i0 .. i0 + n - 1 of stream (seed, stream), shape (n, 3)."""
    f = np.empty((int(n), 3))
    lib().orc_widom_points(seed, stream, i0, int(n), _d(f), nthreads)
    return f


def cell_matrices(a, b, c, alpha, beta, gamma):
    t, inv, v = np.empty(9), np.empty(9), ctypes.c_double()
    lib().orc_cell_matrices(a, b, c, alpha, beta, gamma, _d(t), _d(inv), ctypes.byref(v))
    return t.reshape(3, 3), inv.reshape(3, 3), v.value


def lb_mix(eps_g, sig_g, eps_i, sig_i):
    e, s = ctypes.c_double(), ctypes.c_double()
    lib().orc_lb_mix(eps_g, sig_g, eps_i, sig_i, ctypes.byref(e), ctypes.byref(s))
    return e.value, s.value


def image_atoms(t, atoms, rep):
    atoms = _c(atoms).reshape(-1, 5)
    ni = (2 * rep[0] + 1) * (2 * rep[1] + 1) * (2 * rep[2] + 1)
    out = np.empty((ni * len(atoms), 3))
    lib().orc_image_atoms(_d(_c(t)), _d(atoms), len(atoms), _rep(rep), _d(out))
    return out


def lj_energy(t, atoms, rep, rc2, min_r2, f):
    atoms = _c(atoms).reshape(-1, 5)
    return lib().orc_lj_energy(_d(_c(t)), _d(atoms), len(atoms), _rep(rep), rc2, min_r2,
                               _d(_c(f)))


def lj_grid(t, atoms, rep, rc2, n, sheets=None, nthreads=0):
    """Grid energies; sheets = (begin, end, step) over i (default all)."""
    atoms = _c(atoms).reshape(-1, 5)
    b, e, s = sheets if sheets else (0, n[0], 1)
    ns = len(range(b, e, s))
    out = np.empty((ns, n[1], n[2]))
    lib().orc_lj_grid(_d(_c(t)), _d(atoms), len(atoms), _rep(rep), rc2, n[0], n[1], n[2], b, e,
                      s, _d(out), nthreads)
    return out


def lj_points(t, atoms, rep, rc2, min_r2, f, nthreads=0):
    atoms = _c(atoms).reshape(-1, 5)
    f = _c(f).reshape(-1, 3)
    out = np.empty(len(f))
    lib().orc_lj_points(_d(_c(t)), _d(atoms), len(atoms), _rep(rep), rc2, min_r2, _d(f), len(f),
                        _d(out), nthreads)
    return out


def accessibility_grid(t, atoms, n, nthreads=0):
    """Hard-sphere accessibility (1 accessible / 0 overlap), images [-1, 1]^3."""
    atoms = _c(atoms).reshape(-1, 5)
    out = np.empty(int(n[0]) * int(n[1]) * int(n[2]))
    lib().orc_accessibility_grid(_d(_c(t)), _d(atoms), len(atoms), int(n[0]), int(n[1]),
                                 int(n[2]), _d(out), nthreads)
    return out.reshape(int(n[0]), int(n[1]), int(n[2]))


def count_incut_grid(t, atoms, rep, rc2, n, sheets=None, nthreads=0):
    atoms = _c(atoms).reshape(-1, 5)
    b, e, s = sheets if sheets else (0, n[0], 1)
    return lib().orc_count_incut_grid(_d(_c(t)), _d(atoms), len(atoms), _rep(rep), rc2, n[0],
                                      n[1], n[2], b, e, s, nthreads)


def widom(t, atoms, rep, rc2, temperature, seed, i0, i1, energies=False, nthreads=0):
    atoms = _c(atoms).reshape(-1, 5)
    e = np.empty(i1 - i0) if energies else None
    sb, seb = ctypes.c_double(), ctypes.c_double()
    lib().orc_widom(_d(_c(t)), _d(atoms), len(atoms), _rep(rep), rc2, temperature, seed, i0, i1,
                    _d(e) if energies else None, ctypes.byref(sb), ctypes.byref(seb), nthreads)
    return (sb.value, seb.value, e) if energies else (sb.value, seb.value)


def henry_counts(per_a3: int, volume: float):
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
    lib().orc_henry_counts(per_a3, volume, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    return a.value, b.value, c.value


def henry_result(sum_b, sum_eb, n_total, temperature, density):
    o = [ctypes.c_double() for _ in range(4)]
    lib().orc_henry_result(sum_b, sum_eb, n_total, temperature, density,
                           *[ctypes.byref(x) for x in o])
    return tuple(x.value for x in o)


def find_and_block_pockets(grid, temperature):
    """Returns (npockets, blocked grid copy, segment ids)."""
    g = _c(grid).copy()
    seg = np.empty(g.shape, dtype=np.int32)
    nx, ny, nz = g.shape
    n = lib().orc_find_and_block_pockets(_d(g), nx, ny, nz, temperature,
                                         seg.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return n, g, seg


def wrap(x, z=1.0):
    return lib().orc_wrap(x, z)


def trilinear(grid, xf, yf, zf):
    g = _c(grid)
    nx, ny, nz = g.shape
    return lib().orc_trilinear(_d(g), nx, ny, nz, xf, yf, zf)


def trilinear_points(grid, frac):
    """Interpolated energies at fractional positions frac[n, 3] after the
    wrap x - floor(x) (src/gcmc.cc:74-77, 145-177), one point at a time."""
    g = _c(grid)
    nx, ny, nz = g.shape
    f = _c(frac).reshape(-1, 3)
    L = lib()
    out = np.empty(len(f))
    for i, (x, y, z) in enumerate(f):
        out[i] = L.orc_trilinear(_d(g), nx, ny, nz, L.orc_wrap(x, 1.0), L.orc_wrap(y, 1.0),
                                 L.orc_wrap(z, 1.0))
    return out


def widom_grid(grid, temperature, seed, i0, i1, energies=False, nthreads=0):
    g = _c(grid)
    nx, ny, nz = g.shape
    e = np.empty(i1 - i0) if energies else None
    sb, seb = ctypes.c_double(), ctypes.c_double()
    lib().orc_widom_grid(_d(g), nx, ny, nz, temperature, seed, i0, i1,
                         _d(e) if energies else None, ctypes.byref(sb), ctypes.byref(seb),
                         nthreads)
    return (sb.value, seb.value, e) if energies else (sb.value, seb.value)


def have_ref() -> bool:
    return os.path.exists(REF_HOST)


def replication_factors(t, cutoff: float, twice: bool = False):
    """julia_codes/framework.jl:162-211 restated: per axis start at 1; while the
    distance from the supercell centre (a+b+c)/2 to the plane through the
    origin spanned by the other two original cell vectors is below the cutoff
    (doubled for "twice"), add 1 and double that cell vector."""
    t = np.asarray(t, dtype=np.float64).reshape(3, 3)
    rc = 2.0 * cutoff if twice else cutoff
    v = [t[:, 0].copy(), t[:, 1].copy(), t[:, 2].copy()]
    normals = [np.cross(v[1], v[2]), np.cross(v[0], v[2]), np.cross(v[0], v[1])]
    rep = [1, 1, 1]
    for k in range(3):
        nrm = normals[k]
        while abs(float(np.dot(nrm, 0.5 * (v[0] + v[1] + v[2])))) / float(np.linalg.norm(nrm)) < rc:
            rep[k] += 1
            v[k] = 2.0 * v[k]
    return rep
