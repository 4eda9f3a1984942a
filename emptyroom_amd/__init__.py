"""EmptyRoom on MI355X: the LJ energy grid (writegrid) and Widom Henry
coefficients (henry) as gfx950 HIP kernels behind a C ABI
(include/emptyroom.h, emptyroom_amd/lib/libemptyroom.so), with drop-in
writegrid / henry binaries in emptyroom_amd/bin.

This package only binds the native library (``api``) and generates the
synthetic benchmark frameworks (``synth``)."""
from .api import (ER_FP32, ER_FP64, Batch, Context, EmptyRoomError, Framework,  # noqa: F401
                  HostRun, compute_grid, count_incut_grid, count_incut_widom, henry_result,
                  lib, widom, write_grid_file)
