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
