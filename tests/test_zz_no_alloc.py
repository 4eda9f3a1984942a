"""er_grid_compute_noalloc: the grid written into caller-owned device memory
with no allocation by the library equals er_grid_compute's bit for bit, on
two calls in a row (the second smaller, in FP64), through the C ABI.  The
caller's device buffers come from the HIP runtime the library itself uses."""
import ctypes

import numpy as np
import pytest

from emptyroom_amd import api, synth

pytestmark = pytest.mark.gpu


def hip():
    h = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return h


def noalloc(ctx, fw, n, precision):
    h = hip()
    dt = np.float64 if precision == api.ER_FP64 else np.float32
    out = np.empty(tuple(n), dtype=dt)
    buf = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(buf), out.nbytes) == 0
    try:
        narr = (ctypes.c_int * 3)(*n)
        code = api.lib().er_grid_compute_noalloc(ctx.h, fw.h, narr, precision, buf)
        assert code == 0, api.lib().er_last_error()
        assert h.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), buf, out.nbytes, 2) == 0
    finally:
        h.hipFree(buf)
    return out


def test_noalloc_matches_grid_compute(ctx):
    ctx.tuning = api.default_tuning()
    # a framework of the batching mix at 0.1 A (regions of many tiles), then a
    # small triclinic one at a coarse grid in FP64
    f1 = synth.config3(256)[20]
    f2 = synth.small_framework(5, n_atoms=40, cell=(10.0, 11.0, 12.0, 85.0, 95.0, 100.0))
    cases = [(f1, [60, 70, 80], api.ER_FP32), (f2, [9, 10, 11], api.ER_FP64)]
    for f, n, prec in cases:
        print(f"noalloc {f.name} grid {n} precision {prec}", flush=True)
        a = f.particles()
        fw = api.Framework(ctx, a[:, :3], a[:, 3], a[:, 4], f.t, f.rep, f.rc2)
        ref = api.compute_grid(ctx, fw, n, prec)
        got = noalloc(ctx, fw, n, prec)
        fw.close()
        assert np.array_equal(got, ref)


def test_noalloc_refuses_bad_arguments(ctx):
    f = synth.small_framework(6, n_atoms=10)
    a = f.particles()
    fw = api.Framework(ctx, a[:, :3], a[:, 3], a[:, 4], f.t, f.rep, f.rc2)
    L = api.lib()
    narr = (ctypes.c_int * 3)(1, 5, 5)
    assert L.er_grid_compute_noalloc(ctx.h, fw.h, narr, 0, ctypes.c_void_p(1)) == 1
    narr = (ctypes.c_int * 3)(5, 5, 5)
    assert L.er_grid_compute_noalloc(ctx.h, fw.h, narr, 2, ctypes.c_void_p(1)) == 1
    assert L.er_grid_compute_noalloc(ctx.h, fw.h, narr, 0, None) == 1
    fw.close()
