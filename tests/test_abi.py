"""The C ABI library loads without a GPU and exports every function that
include/emptyroom.h declares (CPU; no compute calls)."""
import ctypes
import os
import re

from emptyroom_amd import api

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(os.path.dirname(HERE), "include", "emptyroom.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(er_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared()
    assert len(names) >= 30
    lib = ctypes.CDLL(api.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_table_covers_header():
    assert sorted(api.EXPORTS) == declared()


def test_version_and_error_without_device():
    L = api.lib()
    assert b"gfx950" in L.er_version()
    # argument errors are reported before any device work
    h = ctypes.c_void_p()
    assert L.er_batch_create(None, None, 0, 0, ctypes.byref(h)) == 1
    assert b"bad arguments" in L.er_last_error()


def test_argument_errors_before_device_work():
    """Null or malformed arguments are refused on the host with ER_E_ARG and
    a message, before any device call (no GPU needed)."""
    L = api.lib()
    out = ctypes.c_void_p()
    res = api.WidomResult()
    t = (ctypes.c_double * 9)(10, 0, 0, 0, 10, 0, 0, 0, 10)
    rep = (ctypes.c_int * 3)(1, 1, 1)
    n = (ctypes.c_int * 3)(4, 4, 4)
    cases = [
        (lambda: L.er_direct_grid(None, None, 0, t, rep, 156.25, n, None, None), b"er_direct_grid"),
        (lambda: L.er_direct_widom(None, None, 0, t, rep, 156.25, 10, 300.0, 1,
                                   ctypes.byref(res), None), b"er_direct_widom"),
        (lambda: L.er_grid_energies(None, None, None, 0, None), b"er_grid_energies"),
        (lambda: L.er_framework_upload(None, None, 0, t, rep, 156.25, ctypes.byref(out)),
         b"er_framework_upload"),
        (lambda: L.er_widom(None, None, 10, 300.0, 1, ctypes.byref(res), None), b"er_widom"),
        (lambda: L.er_grid_tile_for(None, t, n, None, None), b"er_grid_tile_for"),
    ]
    for call, name in cases:
        assert call() == 1
        assert name in L.er_last_error()
