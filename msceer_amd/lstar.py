"""Host-side mirror of the reference's lower-star gradient interface.

``steepest_lstar_grad(values)`` is what steepest_lstar/steepest_lstar.cxx:89-258
computes (max/min labelling, sliding lower-star pairing, ldir pass, flag pass)
and returns the ``.grad`` bytes; ``write_grad`` / ``load_grad`` mirror
DiscreteGradientLabeling::outputToFile / load_from_file
(gi_discrete_gradient_labeling.h:200-253).  Work runs on the GPU through
libmsceer_hip.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import native


def num_cells(X: int, Y: int, Z: int) -> int:
    return (2 * X - 1) * (2 * Y - 1) * (2 * Z - 1)


def steepest_lstar_grad(values: np.ndarray, ctx: native.Context | None = None,
                        skip_ldir: bool = False, timing: bool = False, tiled_layout: bool = False) -> np.ndarray:
    """.grad bytes of a (Z, Y, X) float32 volume, shape (2Z-1, 2Y-1, 2X-1).

    tiled_layout: the tiled program's file (pairs and boundary flag, the
    ascending-manifold dimension bits 0)."""
    v = np.ascontiguousarray(values, dtype=np.float32)
    if v.ndim != 3:
        raise ValueError("values must be a (Z, Y, X) array")
    Z, Y, X = v.shape
    own = ctx is None
    ctx = ctx or native.Context()
    try:
        out = np.empty(num_cells(X, Y, Z), dtype=np.uint8)
        flags = ((native.SKIP_LDIR if skip_ldir else 0) | (native.TIMING if timing else 0) |
                 (native.TILED_LAYOUT if tiled_layout else 0))
        rc = native.lib().msceer_lstar_grad(ctx.handle, v.ctypes.data_as(ctypes.c_void_p), X, Y, Z,
                                            out.ctypes.data_as(ctypes.c_void_p), flags)
        native.check(rc, "msceer_lstar_grad")
    finally:
        if own:
            ctx.close()
    return out.reshape(2 * Z - 1, 2 * Y - 1, 2 * X - 1)


def steepest_lstar_grad_device(ctx: native.Context, values_ptr: int, X: int, Y: int, Z: int,
                               grad_ptr: int, skip_ldir: bool = False, timing: bool = False) -> None:
    """Same on device buffers (pointers from torch tensors or hipMalloc)."""
    flags = native.DEVICE_PTRS | (native.SKIP_LDIR if skip_ldir else 0) | (native.TIMING if timing else 0)
    rc = native.lib().msceer_lstar_grad(ctx.handle, ctypes.c_void_p(values_ptr), X, Y, Z,
                                        ctypes.c_void_p(grad_ptr), flags)
    native.check(rc, "msceer_lstar_grad")


def write_grad(path: str, grad: np.ndarray) -> None:
    np.ascontiguousarray(grad, dtype=np.uint8).tofile(path)


def load_grad(path: str, X: int, Y: int, Z: int) -> np.ndarray:
    g = np.fromfile(path, dtype=np.uint8)
    if g.size != num_cells(X, Y, Z):
        raise ValueError(f"{path}: {g.size} bytes, expected {num_cells(X, Y, Z)}")
    return g.reshape(2 * Z - 1, 2 * Y - 1, 2 * X - 1)


# .grad byte fields (GCC LSB-first bitfield order of GradBitfield)
def field_assigned(g):
    return g & 1


def field_flag(g):
    return (g >> 1) & 1


def field_pair(g):
    return (g >> 2) & 7


def field_ldir(g):
    return g >> 5
