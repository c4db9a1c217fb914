"""Batches of independent volumes (msceer_batch, msceer_batch_grad)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import native
from .lstar import num_cells


def run_batch(volumes, labels: bool = False, streams: int = 0, err: float = 1e-3, gthr: float = 1e-6,
              iters: int = 4096, presimp: float = 0.0, want_grad: bool = True, ctx: native.Context | None = None):
    """For each (Z,Y,X) float32 volume: .grad (and labels when `labels`).

    Returns (grads, labels, ms): lists in input order (None where not
    requested) and the per-volume wall times in milliseconds.
    """
    vols = [np.ascontiguousarray(v, np.float32) for v in volumes]
    n = len(vols)
    arr = (native.Volume * n)()
    grads, labs = [], []
    for i, v in enumerate(vols):
        Z, Y, X = v.shape
        g = np.empty(num_cells(X, Y, Z), np.uint8) if (want_grad or not labels) else None
        l = np.empty(v.size, np.int32) if labels else None
        grads.append(g)
        labs.append(l)
        arr[i].values = v.ctypes.data
        arr[i].nx, arr[i].ny, arr[i].nz = X, Y, Z
        arr[i].grad_out = g.ctypes.data if g is not None else None
        arr[i].labels_out = l.ctypes.data if l is not None else None
    own = ctx is None
    ctx = ctx or native.Context()
    try:
        native.check(native.lib().msceer_batch(ctx.handle, arr, n, streams, err, gthr, iters, presimp, 0),
                     "msceer_batch")
    finally:
        if own:
            ctx.close()
    shaped = [None if l is None else l.reshape(v.shape) for l, v in zip(labs, vols)]
    return grads, shaped, np.array([arr[i].ms for i in range(n)], np.float64)


def volume_table(shapes, values_ptrs, grad_ptrs):
    """A msceer_volume array for .grad-only batches."""
    n = len(shapes)
    arr = (native.Volume * n)()
    for i, ((Z, Y, X), vp, gp) in enumerate(zip(shapes, values_ptrs, grad_ptrs)):
        arr[i].values = vp
        arr[i].nx, arr[i].ny, arr[i].nz = X, Y, Z
        arr[i].grad_out = gp
        arr[i].labels_out = None
    return arr


def run_batch_grad(volumes, max_group_verts: int = 0, timing: bool = False, ctx: native.Context | None = None):
    """Lower-star .grad of each (Z,Y,X) float32 volume through msceer_batch_grad
    (host buffers).  Returns (grads, ms, stats): the .grad arrays in input
    order, per-volume milliseconds (device clock with `timing`, else the
    group's time) and the call's native.BatchStats."""
    vols = [np.ascontiguousarray(v, np.float32) for v in volumes]
    grads = [np.empty(num_cells(v.shape[2], v.shape[1], v.shape[0]), np.uint8) for v in vols]
    arr = volume_table([v.shape for v in vols], [v.ctypes.data for v in vols], [g.ctypes.data for g in grads])
    own = ctx is None
    ctx = ctx or native.Context()
    st = native.BatchStats()
    try:
        native.check(native.lib().msceer_batch_grad(ctx.handle, arr, len(vols), max_group_verts,
                                                    native.TIMING if timing else 0), "msceer_batch_grad")
        native.check(native.lib().msceer_last_batch_stats(ctx.handle, ctypes.byref(st)), "msceer_last_batch_stats")
    finally:
        if own:
            ctx.close()
    return grads, np.array([arr[i].ms for i in range(len(vols))], np.float64), st
