"""ctypes bindings of libemptyroom.so (include/emptyroom.h).

The names follow the reference's own vocabulary: a Framework holds the
FrameworkParticle array of src/datatypes.h:20-28 on the device,
``compute_grid`` replaces the ComputeGridSheet x-sheet loop of
src/writegrid.cu:219-271 and ``widom`` replaces doHenry and its launch
(src/doHenry.h:29-135, src/henry.cu:111-164).  Errors raise EmptyRoomError
with the library's message; nothing here computes anything itself, and the
module fails loudly when the native library has not been built.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libemptyroom.so")
BIN_DIR = os.path.join(_HERE, "bin")

ER_FP32 = 0
ER_FP64 = 1

STATUS = {
    0: "ER_OK", 1: "ER_E_ARG", 2: "ER_E_IO", 3: "ER_E_PARSE", 4: "ER_E_MISSING",
    5: "ER_E_HIP", 6: "ER_E_NOMEM", 7: "ER_E_STATE", 8: "ER_E_CAPACITY", 9: "ER_E_NODEVICE",
}


class EmptyRoomError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{STATUS.get(code, code)}: {message}")
        self.code = code
        self.message = message


class Particle(ctypes.Structure):
    _fields_ = [("x_f", ctypes.c_double), ("y_f", ctypes.c_double), ("z_f", ctypes.c_double),
                ("eps_with_adsorbate", ctypes.c_double),
                ("sig_with_adsorbate_squared", ctypes.c_double)]


class WidomResult(ctypes.Structure):
    _fields_ = [("sum_b", ctypes.c_double), ("sum_eb", ctypes.c_double), ("n", ctypes.c_uint64)]


class Tuning(ctypes.Structure):
    _fields_ = [("points_per_lane", ctypes.c_int), ("lds_records", ctypes.c_int),
                ("tile", ctypes.c_int * 3), ("widom_per_bin", ctypes.c_int),
                ("s2_switch", ctypes.c_double), ("region_len", ctypes.c_double),
                ("class_spacing", ctypes.c_double * 2), ("class_tile", (ctypes.c_int * 3) * 3)]

    def to_dict(self) -> dict:
        return dict(points_per_lane=self.points_per_lane, lds_records=self.lds_records,
                    tile=list(self.tile), widom_per_bin=self.widom_per_bin,
                    s2_switch=self.s2_switch, region_len=self.region_len,
                    class_spacing=list(self.class_spacing),
                    class_tile=[list(r) for r in self.class_tile])

    @classmethod
    def from_dict(cls, d: dict, base: "Tuning" = None) -> "Tuning":
        t = cls()
        if base is not None:
            ctypes.memmove(ctypes.byref(t), ctypes.byref(base), ctypes.sizeof(cls))
        for k, v in d.items():
            if k in ("tile", "class_spacing"):
                for i, x in enumerate(v):
                    getattr(t, k)[i] = x
            elif k == "class_tile":
                for c, row in enumerate(v):
                    for i, x in enumerate(row):
                        t.class_tile[c][i] = x
            elif k in dict(cls._fields_):
                setattr(t, k, v)
            else:
                raise KeyError(f"unknown tuning field {k}")
        return t


class Task(ctypes.Structure):
    _fields_ = [("fw", ctypes.c_void_p), ("grid_n", ctypes.c_int * 3),
                ("precision", ctypes.c_int), ("n_insertions", ctypes.c_uint64),
                ("temperature", ctypes.c_double), ("seed", ctypes.c_uint64)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("streams", ctypes.c_int), ("items_per_launch", ctypes.c_int64),
                ("max_queue", ctypes.c_int), ("cache_bytes", ctypes.c_int64),
                ("fail_launch", ctypes.c_int)]


class Request(ctypes.Structure):
    _fields_ = [("atoms", ctypes.POINTER(Particle)), ("n_atoms", ctypes.c_int),
                ("t_matrix", ctypes.c_double * 9), ("rep", ctypes.c_int * 3),
                ("rc2", ctypes.c_double), ("grid_n", ctypes.c_int * 3),
                ("precision", ctypes.c_int), ("n_insertions", ctypes.c_uint64),
                ("temperature", ctypes.c_double), ("seed", ctypes.c_uint64)]


class RequestStatus(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int), ("status", ctypes.c_int), ("t_submit", ctypes.c_double),
                ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double),
                ("widom", WidomResult)]


RQ_QUEUED, RQ_RUNNING, RQ_DONE, RQ_FAILED, RQ_REJECTED = range(5)


class HostInfo(ctypes.Structure):
    _fields_ = [("a", ctypes.c_double), ("b", ctypes.c_double), ("c", ctypes.c_double),
                ("alpha", ctypes.c_double), ("beta", ctypes.c_double), ("gamma", ctypes.c_double),
                ("n_atoms", ctypes.c_int), ("density", ctypes.c_double),
                ("volume", ctypes.c_double), ("t_matrix", ctypes.c_double * 9),
                ("inv_t_matrix", ctypes.c_double * 9), ("rep", ctypes.c_int * 3),
                ("r_cutoff_squared", ctypes.c_double), ("temperature", ctypes.c_double),
                ("epsilon_guest", ctypes.c_double), ("sigma_guest", ctypes.c_double),
                ("adsorbate_mw", ctypes.c_double), ("grid_n", ctypes.c_int * 3),
                ("grid_resolution", ctypes.c_double), ("grid_precision", ctypes.c_int),
                ("insertions_per_a3", ctypes.c_int), ("numinsertions", ctypes.c_int),
                ("insertions_per_thread", ctypes.c_int), ("total_insertions", ctypes.c_int64),
                ("seed", ctypes.c_uint64)]


# every symbol include/emptyroom.h declares
EXPORTS = [
    "er_last_error", "er_version", "er_context_create", "er_context_destroy",
    "er_context_get_tuning", "er_context_set_tuning", "er_context_memory",
    "er_framework_upload", "er_framework_release", "er_grid_compute", "er_grid_download",
    "er_grid_info", "er_grid_release", "er_widom", "er_count_incut_grid",
    "er_count_incut_grid_map", "er_count_incut_widom", "er_widom_point",
    "er_grid_compute_noalloc", "er_henry_result", "er_batch_create", "er_batch_run",
    "er_batch_wait", "er_batch_result", "er_batch_release", "er_batch_set_timing",
    "er_batch_kernel_ms", "er_grid_upload", "er_pockets_label", "er_pockets_label_batch",
    "er_widom_grid",
    "er_server_create", "er_server_submit", "er_server_poll", "er_server_fetch_grid",
    "er_server_release", "er_server_stats", "er_server_cache_stats", "er_server_destroy",
    "er_host_setup",
    "er_host_run_info", "er_host_run_particles", "er_host_run_label",
    "er_host_run_write_settings", "er_host_run_release", "er_write_grid_file",
    "er_replication_factors", "er_write_uc_file", "er_accessibility_grid",
    "er_grid_file_info", "er_read_grid_file", "er_write_grid_binary", "er_grid_load",
    "er_tuning_defaults", "er_grid_tile_for", "er_grid_energies", "er_direct_grid",
    "er_direct_widom",
]

_lib: Optional[ctypes.CDLL] = None

_P = ctypes.c_void_p
_SIGS = {
    "er_last_error": (ctypes.c_char_p, []),
    "er_version": (ctypes.c_char_p, []),
    "er_context_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(_P)]),
    "er_context_destroy": (ctypes.c_int, [_P]),
    "er_context_get_tuning": (ctypes.c_int, [_P, ctypes.POINTER(Tuning)]),
    "er_context_set_tuning": (ctypes.c_int, [_P, ctypes.POINTER(Tuning)]),
    "er_context_memory": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_size_t),
                                         ctypes.POINTER(ctypes.c_size_t)]),
    "er_framework_upload": (ctypes.c_int, [_P, ctypes.POINTER(Particle), ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_int), ctypes.c_double,
                                           ctypes.POINTER(_P)]),
    "er_framework_release": (ctypes.c_int, [_P]),
    "er_grid_compute": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_int), ctypes.c_int, _P,
                                       ctypes.POINTER(_P)]),
    "er_grid_download": (ctypes.c_int, [_P, _P]),
    "er_grid_info": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int),
                                    ctypes.POINTER(ctypes.c_int)]),
    "er_grid_release": (ctypes.c_int, [_P]),
    "er_widom": (ctypes.c_int, [_P, _P, ctypes.c_uint64, ctypes.c_double, ctypes.c_uint64,
                                ctypes.POINTER(WidomResult), _P]),
    "er_count_incut_grid": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_int),
                                           ctypes.POINTER(ctypes.c_uint64)]),
    "er_count_incut_grid_map": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_int),
                                               ctypes.POINTER(ctypes.c_uint64), _P]),
    "er_grid_compute_noalloc": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_int),
                                               ctypes.c_int, _P]),
    "er_count_incut_widom": (ctypes.c_int, [_P, _P, ctypes.c_uint64, ctypes.c_uint64,
                                            ctypes.POINTER(ctypes.c_uint64)]),
    "er_widom_point": (None, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64,
                              ctypes.POINTER(ctypes.c_double)]),
    "er_henry_result": (None, [ctypes.c_double] * 5 + [ctypes.POINTER(ctypes.c_double)] * 4),
    "er_batch_create": (ctypes.c_int, [_P, ctypes.POINTER(Task), ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(_P)]),
    "er_batch_run": (ctypes.c_int, [_P]),
    "er_batch_wait": (ctypes.c_int, [_P]),
    "er_batch_result": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(WidomResult),
                                       ctypes.POINTER(_P)]),
    "er_batch_release": (ctypes.c_int, [_P]),
    "er_batch_set_timing": (ctypes.c_int, [_P, ctypes.c_int]),
    "er_batch_kernel_ms": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_int)]),
    "er_grid_upload": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(_P)]),
    "er_pockets_label": (ctypes.c_int, [_P, _P, ctypes.c_double, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_int), _P]),
    "er_pockets_label_batch": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.c_int,
                                              ctypes.c_double, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_int)]),
    "er_widom_grid": (ctypes.c_int, [_P, _P, ctypes.c_uint64, ctypes.c_double, ctypes.c_uint64,
                                     ctypes.POINTER(WidomResult), _P]),
    "er_server_create": (ctypes.c_int, [_P, ctypes.POINTER(ServerConfig), ctypes.POINTER(_P)]),
    "er_server_submit": (ctypes.c_int, [_P, ctypes.POINTER(Request),
                                        ctypes.POINTER(ctypes.c_uint64)]),
    "er_server_poll": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_double,
                                      ctypes.POINTER(RequestStatus)]),
    "er_server_fetch_grid": (ctypes.c_int, [_P, ctypes.c_uint64, _P]),
    "er_server_release": (ctypes.c_int, [_P, ctypes.c_uint64]),
    "er_server_stats": (ctypes.c_int, [_P] + [ctypes.POINTER(ctypes.c_int64)] * 4),
    "er_server_cache_stats": (ctypes.c_int, [_P] + [ctypes.POINTER(ctypes.c_int64)] * 5),
    "er_server_destroy": (ctypes.c_int, [_P]),
    "er_host_setup": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                     ctypes.c_int, ctypes.POINTER(_P)]),
    "er_host_run_info": (ctypes.c_int, [_P, ctypes.POINTER(HostInfo)]),
    "er_host_run_particles": (ctypes.c_int, [_P, ctypes.POINTER(Particle), ctypes.c_int]),
    "er_host_run_label": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]),
    "er_host_run_write_settings": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "er_host_run_release": (None, [_P]),
    "er_write_grid_file": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _P, _P,
                                          ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_double), ctypes.c_int]),
    "er_grid_file_info": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]),
    "er_read_grid_file": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), _P]),
    "er_write_grid_binary": (ctypes.c_int, [ctypes.c_char_p, _P, _P,
                                            ctypes.POINTER(ctypes.c_int), _P]),
    "er_grid_load": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.POINTER(_P)]),
    "er_accessibility_grid": (ctypes.c_int, [_P, _P, ctypes.c_int, _P,
                                             ctypes.POINTER(ctypes.c_int), _P]),
    "er_direct_grid": (ctypes.c_int, [_P, _P, ctypes.c_int, _P, ctypes.POINTER(ctypes.c_int),
                                      ctypes.c_double, ctypes.POINTER(ctypes.c_int), _P,
                                      ctypes.POINTER(_P)]),
    "er_direct_widom": (ctypes.c_int, [_P, _P, ctypes.c_int, _P, ctypes.POINTER(ctypes.c_int),
                                       ctypes.c_double, ctypes.c_uint64, ctypes.c_double,
                                       ctypes.c_uint64, ctypes.POINTER(WidomResult), _P]),
    "er_tuning_defaults": (ctypes.c_int, [ctypes.POINTER(Tuning)]),
    "er_grid_energies": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_double), ctypes.c_uint64,
                                        ctypes.POINTER(ctypes.c_double)]),
    "er_grid_tile_for": (ctypes.c_int, [ctypes.POINTER(Tuning), ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_int)]),
    "er_replication_factors": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_int)]),
    "er_write_uc_file": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_double,
                                        ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
}


def lib() -> ctypes.CDLL:
    """Load libemptyroom.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(make -C emptyroom_amd)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(code: int) -> None:
    if code != 0:
        raise EmptyRoomError(code, lib().er_last_error().decode(errors="replace"))


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _iarr(v: Sequence[int]):
    return (ctypes.c_int * len(v))(*[int(x) for x in v])


def particles_array(frac: np.ndarray, eps: np.ndarray, sig2: np.ndarray):
    """Pack per-atom (x_f, y_f, z_f, eps_ij, sig2_ij) into a FrameworkParticle array."""
    n = len(frac)
    arr = (Particle * max(n, 1))()
    buf = np.frombuffer(arr, dtype=np.float64).reshape(max(n, 1), 5)
    if n:
        buf[:n, 0:3] = frac
        buf[:n, 3] = eps
        buf[:n, 4] = sig2
    return arr, n


class Context:
    def __init__(self, device: int = 0, hbm_cap: int = 0):
        h = _P()
        _check(lib().er_context_create(device, hbm_cap, ctypes.byref(h)))
        self.h = h

    def close(self) -> None:
        if self.h:
            lib().er_context_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def tuning(self) -> Tuning:
        t = Tuning()
        _check(lib().er_context_get_tuning(self.h, ctypes.byref(t)))
        return t

    @tuning.setter
    def tuning(self, t: Tuning) -> None:
        _check(lib().er_context_set_tuning(self.h, ctypes.byref(t)))

    def memory(self):
        a, b = ctypes.c_size_t(), ctypes.c_size_t()
        _check(lib().er_context_memory(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value


def default_tuning() -> Tuning:
    """The library's default kernel knobs (host only, no device needed)."""
    t = Tuning()
    _check(lib().er_tuning_defaults(ctypes.byref(t)))
    return t


def grid_tile_for(tuning: Tuning, t_matrix, n: Sequence[int]):
    """(tile, size class) that `tuning` gives a grid of n points on t_matrix."""
    tm = np.ascontiguousarray(np.asarray(t_matrix, dtype=np.float64).reshape(9))
    nn = (ctypes.c_int * 3)(*[int(x) for x in n])
    tile = (ctypes.c_int * 3)()
    cls = ctypes.c_int()
    _check(lib().er_grid_tile_for(ctypes.byref(tuning), _dptr(tm), nn, tile, ctypes.byref(cls)))
    return tuple(tile), cls.value


class Framework:
    """A framework resident in HBM (FrameworkParticle array + cell)."""

    def __init__(self, ctx: Context, frac: np.ndarray, eps: np.ndarray, sig2: np.ndarray,
                 t_matrix: np.ndarray, rep: Sequence[int], rc2: float):
        arr, n = particles_array(np.asarray(frac, dtype=np.float64),
                                 np.asarray(eps, dtype=np.float64),
                                 np.asarray(sig2, dtype=np.float64))
        t = np.ascontiguousarray(np.asarray(t_matrix, dtype=np.float64).reshape(9))
        h = _P()
        _check(lib().er_framework_upload(ctx.h, arr, n, _dptr(t), _iarr(rep), float(rc2),
                                         ctypes.byref(h)))
        self.h = h
        self.ctx = ctx
        self.n_atoms = n

    def close(self) -> None:
        if self.h:
            lib().er_framework_release(self.h)
            self.h = None


def compute_grid(ctx: Context, fw: Framework, n: Sequence[int], precision: int = ER_FP32
                 ) -> np.ndarray:
    """Energy (K) at every grid point, shape (nx, ny, nz)."""
    dt = np.float64 if precision == ER_FP64 else np.float32
    out = np.empty(int(n[0]) * int(n[1]) * int(n[2]), dtype=dt)
    _check(lib().er_grid_compute(ctx.h, fw.h, _iarr(n), precision,
                                 out.ctypes.data_as(ctypes.c_void_p), None))
    return out.reshape(int(n[0]), int(n[1]), int(n[2]))


def widom(ctx: Context, fw: Framework, n: int, temperature: float, seed: int,
          energies: bool = False):
    """Widom sums (sum_b, sum_eb, n) and optionally E per insertion."""
    res = WidomResult()
    e = np.empty(int(n), dtype=np.float64) if energies else None
    _check(lib().er_widom(ctx.h, fw.h, int(n), float(temperature), int(seed), ctypes.byref(res),
                          e.ctypes.data_as(ctypes.c_void_p) if energies else None))
    return (res.sum_b, res.sum_eb, res.n, e) if energies else (res.sum_b, res.sum_eb, res.n)


def count_incut_grid(ctx: Context, fw: Framework, n: Sequence[int]) -> int:
    v = ctypes.c_uint64()
    _check(lib().er_count_incut_grid(ctx.h, fw.h, _iarr(n), ctypes.byref(v)))
    return v.value


def count_incut_grid_map(ctx: Context, fw: Framework, n: Sequence[int]):
    """(total, per-point uint32 counts shaped like the grid)."""
    v = ctypes.c_uint64()
    out = np.empty(tuple(int(x) for x in n), dtype=np.uint32)
    _check(lib().er_count_incut_grid_map(ctx.h, fw.h, _iarr(n), ctypes.byref(v),
                                         out.ctypes.data_as(ctypes.c_void_p)))
    return v.value, out


def count_incut_widom(ctx: Context, fw: Framework, n: int, seed: int) -> int:
    v = ctypes.c_uint64()
    _check(lib().er_count_incut_widom(ctx.h, fw.h, int(n), int(seed), ctypes.byref(v)))
    return v.value


def widom_point(seed: int, stream: int, i: int) -> np.ndarray:
    """Fractional coordinates of insertion i of stream (seed, stream)."""
    f = np.empty(3)
    lib().er_widom_point(int(seed), int(stream), int(i), _dptr(f))
    return f


def henry_result(sum_b: float, sum_eb: float, n_total: float, temperature: float,
                 density: float):
    """(K_H mol/(kg Pa), K_H mol/(m^3 Pa), <E> K, <E> kJ/mol), src/henry.cu:166-174."""
    o = [ctypes.c_double() for _ in range(4)]
    lib().er_henry_result(sum_b, sum_eb, float(n_total), temperature, density,
                          *[ctypes.byref(x) for x in o])
    return tuple(x.value for x in o)


class Batch:
    """Many (framework, grid, Widom) tasks in one persistent launch / hipGraph."""

    def __init__(self, ctx: Context, tasks: Sequence[dict], use_graph: bool = True):
        arr = (Task * max(len(tasks), 1))()
        for i, t in enumerate(tasks):
            arr[i].fw = t["fw"].h
            g = t.get("grid_n") or (0, 0, 0)
            for d in range(3):
                arr[i].grid_n[d] = int(g[d])
            arr[i].precision = int(t.get("precision", ER_FP32))
            arr[i].n_insertions = int(t.get("n_insertions", 0))
            arr[i].temperature = float(t.get("temperature", 298.0))
            arr[i].seed = int(t.get("seed", 0))
        h = _P()
        _check(lib().er_batch_create(ctx.h, arr, len(tasks), int(use_graph), ctypes.byref(h)))
        self.h = h
        self.tasks = list(tasks)

    def run(self) -> None:
        _check(lib().er_batch_run(self.h))

    def wait(self) -> None:
        _check(lib().er_batch_wait(self.h))

    def set_timing(self, on: bool = True) -> None:
        _check(lib().er_batch_set_timing(self.h, int(on)))

    def kernel_ms(self):
        """(milliseconds, launches) of the tiled-kernel launches of the last eager run."""
        ms, n = ctypes.c_double(), ctypes.c_int()
        _check(lib().er_batch_kernel_ms(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def widom(self, i: int):
        r = WidomResult()
        _check(lib().er_batch_result(self.h, i, ctypes.byref(r), None))
        return r.sum_b, r.sum_eb, r.n

    def grid(self, i: int) -> Optional[np.ndarray]:
        g = _P()
        _check(lib().er_batch_result(self.h, i, None, ctypes.byref(g)))
        if not g:
            return None
        n = (ctypes.c_int * 3)()
        prec = ctypes.c_int()
        _check(lib().er_grid_info(g, n, ctypes.byref(prec)))
        dt = np.float64 if prec.value == ER_FP64 else np.float32
        out = np.empty(n[0] * n[1] * n[2], dtype=dt)
        _check(lib().er_grid_download(g, out.ctypes.data_as(ctypes.c_void_p)))
        return out.reshape(n[0], n[1], n[2])

    def close(self) -> None:
        if self.h:
            lib().er_batch_release(self.h)
            self.h = None


class Grid:
    """A resident FP32/FP64 grid (from er_grid_upload or er_grid_compute)."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def upload(cls, ctx: Context, grid: np.ndarray) -> "Grid":
        g = np.ascontiguousarray(grid, dtype=np.float32)
        h = _P()
        _check(lib().er_grid_upload(ctx.h, g.ctypes.data_as(ctypes.c_void_p), _iarr(g.shape),
                                    ctypes.byref(h)))
        return cls(h)

    @classmethod
    def compute(cls, ctx: Context, fw: Framework, n: Sequence[int], precision: int = ER_FP32):
        h = _P()
        _check(lib().er_grid_compute(ctx.h, fw.h, _iarr(n), precision, None, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def load(cls, ctx: Context, path: str) -> "Grid":
        """A text or binary grid file as a resident FP32 grid."""
        h = _P()
        _check(lib().er_grid_load(ctx.h, path.encode(), ctypes.byref(h)))
        return cls(h)

    def download(self) -> np.ndarray:
        n = (ctypes.c_int * 3)()
        prec = ctypes.c_int()
        _check(lib().er_grid_info(self.h, n, ctypes.byref(prec)))
        out = np.empty(n[0] * n[1] * n[2], dtype=np.float64 if prec.value else np.float32)
        _check(lib().er_grid_download(self.h, out.ctypes.data_as(ctypes.c_void_p)))
        return out.reshape(n[0], n[1], n[2])

    def close(self) -> None:
        if self.h:
            lib().er_grid_release(self.h)
            self.h = None


def pockets_label(ctx: Context, grid: Grid, temperature: float, block: bool = True,
                  segments: bool = False):
    """(n_pockets, segment ids or None); blocks pocket voxels in place if asked."""
    npk = ctypes.c_int()
    seg = None
    if segments:
        n = (ctypes.c_int * 3)()
        _check(lib().er_grid_info(grid.h, n, None))
        seg = np.empty((n[0], n[1], n[2]), dtype=np.int32)
    _check(lib().er_pockets_label(ctx.h, grid.h, float(temperature), int(block),
                                  ctypes.byref(npk),
                                  seg.ctypes.data_as(ctypes.c_void_p) if segments else None))
    return npk.value, seg


def accessibility_grid(ctx: Context, frac: np.ndarray, eps: np.ndarray, sig2: np.ndarray,
                       t_matrix: np.ndarray, n: Sequence[int]) -> np.ndarray:
    """Hard-sphere accessibility grid (1 accessible / 0 overlap) as float64."""
    arr, na = particles_array(np.asarray(frac, dtype=np.float64),
                              np.asarray(eps, dtype=np.float64),
                              np.asarray(sig2, dtype=np.float64))
    t = np.ascontiguousarray(np.asarray(t_matrix, dtype=np.float64).reshape(9))
    out = np.empty(int(n[0]) * int(n[1]) * int(n[2]), dtype=np.float64)
    _check(lib().er_accessibility_grid(ctx.h, ctypes.cast(arr, ctypes.c_void_p), na,
                                       t.ctypes.data_as(ctypes.c_void_p), _iarr(n),
                                       out.ctypes.data_as(ctypes.c_void_p)))
    return out.reshape(int(n[0]), int(n[1]), int(n[2]))


def _direct_args(frac, eps, sig2, t_matrix, rep):
    arr, na = particles_array(np.asarray(frac, dtype=np.float64),
                              np.asarray(eps, dtype=np.float64),
                              np.asarray(sig2, dtype=np.float64))
    t = np.ascontiguousarray(np.asarray(t_matrix, dtype=np.float64).reshape(9))
    return arr, na, t, _iarr(rep)


def direct_grid(ctx: Context, frac: np.ndarray, eps: np.ndarray, sig2: np.ndarray,
                t_matrix: np.ndarray, rep: Sequence[int], rc2: float, n: Sequence[int],
                download: bool = True):
    """The reference's direct method on the GPU (FP64, every triple): the grid
    as float64 (download=True) or a resident Grid."""
    arr, na, t, r = _direct_args(frac, eps, sig2, t_matrix, rep)
    out = np.empty(int(n[0]) * int(n[1]) * int(n[2]), dtype=np.float64) if download else None
    g = _P()
    _check(lib().er_direct_grid(ctx.h, ctypes.cast(arr, ctypes.c_void_p), na,
                                t.ctypes.data_as(ctypes.c_void_p), r, float(rc2), _iarr(n),
                                out.ctypes.data_as(ctypes.c_void_p) if download else None,
                                None if download else ctypes.byref(g)))
    return out.reshape(int(n[0]), int(n[1]), int(n[2])) if download else Grid(g)


def direct_widom(ctx: Context, frac: np.ndarray, eps: np.ndarray, sig2: np.ndarray,
                 t_matrix: np.ndarray, rep: Sequence[int], rc2: float, n: int,
                 temperature: float, seed: int, energies: bool = False):
    """Widom sums (sum_b, sum_eb, n) by the direct method at the points of
    widom(); optionally E per insertion."""
    arr, na, t, r = _direct_args(frac, eps, sig2, t_matrix, rep)
    res = WidomResult()
    e = np.empty(int(n), dtype=np.float64) if energies else None
    _check(lib().er_direct_widom(ctx.h, ctypes.cast(arr, ctypes.c_void_p), na,
                                 t.ctypes.data_as(ctypes.c_void_p), r, float(rc2), int(n),
                                 float(temperature), int(seed), ctypes.byref(res),
                                 e.ctypes.data_as(ctypes.c_void_p) if energies else None))
    return (res.sum_b, res.sum_eb, res.n, e) if energies else (res.sum_b, res.sum_eb, res.n)


def pockets_label_batch(ctx: Context, grids: Sequence[Grid], temperature: float,
                        block: bool = True):
    """Pocket counts of many grids labelled (and blocked) together."""
    arr = (_P * max(len(grids), 1))(*[g.h for g in grids])
    out = (ctypes.c_int * max(len(grids), 1))()
    _check(lib().er_pockets_label_batch(ctx.h, arr, len(grids), float(temperature), int(block),
                                        out))
    return [out[i] for i in range(len(grids))]


def grid_energies(ctx: Context, grid: Grid, frac: np.ndarray) -> np.ndarray:
    """Guest-framework energies (K) at fractional positions frac[n, 3] (any
    range; wrapped into [0, 1)) by trilinear interpolation of the grid."""
    f = np.ascontiguousarray(np.asarray(frac, dtype=np.float64).reshape(-1, 3))
    out = np.empty(len(f), dtype=np.float64)
    _check(lib().er_grid_energies(ctx.h, grid.h, _dptr(f), len(f), _dptr(out)))
    return out


def widom_grid(ctx: Context, grid: Grid, n: int, temperature: float, seed: int,
               energies: bool = False):
    res = WidomResult()
    e = np.empty(int(n), dtype=np.float64) if energies else None
    _check(lib().er_widom_grid(ctx.h, grid.h, int(n), float(temperature), int(seed),
                               ctypes.byref(res),
                               e.ctypes.data_as(ctypes.c_void_p) if energies else None))
    return (res.sum_b, res.sum_eb, res.n, e) if energies else (res.sum_b, res.sum_eb, res.n)


class Server:
    """Continuous-batching server on one context (er_server_*)."""

    def __init__(self, ctx: Context, streams: int = 3, items_per_launch: int = 32768,
                 max_queue: int = 4096, cache_bytes: int = -1, fail_launch: int = 0):
        cfg = ServerConfig(streams, items_per_launch, max_queue, cache_bytes, fail_launch)
        h = _P()
        _check(lib().er_server_create(ctx.h, ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self._keep = {}

    def submit(self, frac, eps, sig2, t_matrix, rep, rc2, grid_n=(0, 0, 0),
               precision=ER_FP32, n_insertions=0, temperature=298.0, seed=0):
        """Returns (status code, ticket); codes 6 and 8 are rejections."""
        arr, n = particles_array(np.asarray(frac, dtype=np.float64),
                                 np.asarray(eps, dtype=np.float64),
                                 np.asarray(sig2, dtype=np.float64))
        rq = Request()
        rq.atoms = ctypes.cast(arr, ctypes.POINTER(Particle))
        rq.n_atoms = n
        t = np.asarray(t_matrix, dtype=np.float64).reshape(9)
        for i in range(9):
            rq.t_matrix[i] = t[i]
        for d in range(3):
            rq.rep[d] = int(rep[d])
            rq.grid_n[d] = int(grid_n[d])
        rq.rc2 = float(rc2)
        rq.precision = int(precision)
        rq.n_insertions = int(n_insertions)
        rq.temperature = float(temperature)
        rq.seed = int(seed)
        tk = ctypes.c_uint64()
        code = lib().er_server_submit(self.h, ctypes.byref(rq), ctypes.byref(tk))
        if code not in (0, 6, 8):
            _check(code)
        return code, tk.value

    def poll(self, ticket: int, wait_ms: float = 0.0) -> RequestStatus:
        st = RequestStatus()
        _check(lib().er_server_poll(self.h, int(ticket), float(wait_ms), ctypes.byref(st)))
        return st

    def fetch_grid(self, ticket: int, shape, precision=ER_FP32) -> np.ndarray:
        out = np.empty(tuple(shape), dtype=np.float64 if precision == ER_FP64 else np.float32)
        _check(lib().er_server_fetch_grid(self.h, int(ticket),
                                          out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def release(self, ticket: int) -> None:
        _check(lib().er_server_release(self.h, int(ticket)))

    def stats(self):
        v = [ctypes.c_int64() for _ in range(4)]
        _check(lib().er_server_stats(self.h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("submitted", "rejected", "completed", "launches"), [x.value for x in v]))

    def cache_stats(self):
        v = [ctypes.c_int64() for _ in range(5)]
        _check(lib().er_server_cache_stats(self.h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("hits", "misses", "evictions", "resident_frameworks",
                         "resident_bytes"), [x.value for x in v]))

    def close(self) -> None:
        if self.h:
            lib().er_server_destroy(self.h)
            self.h = None


class HostRun:
    """Host-side setup of a writegrid (mode 0) or henry (mode 1) run."""

    def __init__(self, root: str, structure: str, adsorbate: str, mode: int):
        h = _P()
        _check(lib().er_host_setup(root.encode(), structure.encode(), adsorbate.encode(), mode,
                                   ctypes.byref(h)))
        self.h = h
        self.mode = mode

    def info(self) -> HostInfo:
        i = HostInfo()
        _check(lib().er_host_run_info(self.h, ctypes.byref(i)))
        return i

    def particles(self) -> np.ndarray:
        n = self.info().n_atoms
        arr = (Particle * max(n, 1))()
        _check(lib().er_host_run_particles(self.h, arr, max(n, 1)))
        return np.frombuffer(arr, dtype=np.float64).reshape(max(n, 1), 5)[:n].copy()

    def labels(self):
        n = self.info().n_atoms
        buf = ctypes.create_string_buffer(64)
        out = []
        for i in range(n):
            _check(lib().er_host_run_label(self.h, i, buf, 64))
            out.append(buf.value.decode())
        return out

    def write_settings(self, path: str) -> None:
        _check(lib().er_host_run_write_settings(self.h, path.encode()))

    def close(self) -> None:
        if self.h:
            lib().er_host_run_release(self.h)
            self.h = None


def read_grid_file(path: str) -> np.ndarray:
    """Text (reference format) or binary grid file as an FP64 array (N_x, N_y, N_z)."""
    n = (ctypes.c_int * 3)()
    _check(lib().er_grid_file_info(path.encode(), n))
    out = np.empty(n[0] * n[1] * n[2], dtype=np.float64)
    _check(lib().er_read_grid_file(path.encode(), n, out.ctypes.data_as(ctypes.c_void_p)))
    return out.reshape(n[0], n[1], n[2])


def write_grid_binary(path: str, grid: np.ndarray, t_matrix: np.ndarray) -> None:
    g = np.ascontiguousarray(grid)
    if g.dtype not in (np.float32, np.float64):
        g = g.astype(np.float64)
    t = np.ascontiguousarray(np.asarray(t_matrix, dtype=np.float64).reshape(9))
    p = g.ctypes.data_as(ctypes.c_void_p)
    _check(lib().er_write_grid_binary(path.encode(), p if g.dtype == np.float64 else None,
                                      p if g.dtype == np.float32 else None, _iarr(g.shape),
                                      t.ctypes.data_as(ctypes.c_void_p)))


def write_grid_file(path: str, fmt: str, grid: np.ndarray, t_matrix: np.ndarray,
                    threads: int = 4) -> None:
    g = np.ascontiguousarray(grid)
    n = _iarr(g.shape)
    t = np.ascontiguousarray(np.asarray(t_matrix, dtype=np.float64).reshape(9))
    p = g.ctypes.data_as(ctypes.c_void_p)
    e64, e32 = (p, None) if g.dtype == np.float64 else (None, p)
    _check(lib().er_write_grid_file(path.encode(), fmt.encode(), e64, e32, n, _dptr(t), threads))
