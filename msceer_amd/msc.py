"""The .grad consumer of extractmsc, before any simplification
(msceer_msc_base; extractmsc/extractmsc.cxx:467-556): Morse-Smale complex
nodes = critical cells in cell-id order, the base basins (TESTBASINS.raw) and
the base mountains (TESTMOUNTS.raw).  Runs on the GPU; no CPU path."""
from __future__ import annotations

import ctypes

import numpy as np

from . import native


def msc_base(grad: np.ndarray, shape_zyx, ctx: native.Context | None = None):
    """(basins, mounts, counts) of a .grad of a (Z, Y, X) volume: int32 node
    ids per vertex (-1 where none) and the critical cells by dimension."""
    Z, Y, X = shape_zyx
    g = np.ascontiguousarray(grad, np.uint8).ravel()
    if g.size != (2 * X - 1) * (2 * Y - 1) * (2 * Z - 1):
        raise ValueError("grad does not match the shape")
    basins = np.empty(X * Y * Z, np.int32)
    mounts = np.empty(X * Y * Z, np.int32)
    counts = np.zeros(4, np.int64)
    own = ctx is None
    ctx = ctx or native.Context()
    try:
        native.check(native.lib().msceer_msc_base(ctx.handle, g.ctypes.data_as(ctypes.c_void_p), X, Y, Z,
                                                  basins.ctypes.data_as(ctypes.c_void_p),
                                                  mounts.ctypes.data_as(ctypes.c_void_p),
                                                  counts.ctypes.data_as(ctypes.c_void_p), 0), "msceer_msc_base")
    finally:
        if own:
            ctx.close()
    return basins.reshape(Z, Y, X), mounts.reshape(Z, Y, X), counts
