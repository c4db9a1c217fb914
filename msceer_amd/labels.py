"""Host-side mirror of the reference's accurate-labelling interface.

``ondemand_asc`` is the needASC2 branch of ondemandaccurate
(ondemandaccurate/ondemandaccurate.cxx:728-1061): accurate ascending 3-manifold
labels of every vertex and the conforming ``.grad``.  ``write_labels`` writes
the int32-per-vertex label file of accurate3dmorselabeling
(accurate3dmorselabeling.cxx:333-335, ``<file>._int_XxYxZ.raw``).  Work runs on
the GPU through libmsceer_hip.so.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import native
from .lstar import num_cells


def ondemand_asc(values: np.ndarray, err_thr: float = 1e-3, grad_thr: float = 1e-6, max_iters: int = 4096,
                 presimp: float = 0.0, mask: np.ndarray | None = None, want_grad: bool = True,
                 ctx: native.Context | None = None, stop: str | None = None, fp32_first: bool = False):
    """(grad or None, labels (Z,Y,X) int32, stats dict) of a (Z,Y,X) float32 volume.

    mask: optional per-vertex uint8; only nonzero vertices are integrated,
    the others take the label of the minimum their discrete-gradient V-path
    reaches.  stop: None, "regions" or "integrated" (earlier stage labels).
    fp32_first: re-integrate in FP32 and in FP64 only where the FP32 trace is
    in doubt (MSCEER_FP32_FIRST).
    """
    v = np.ascontiguousarray(values, dtype=np.float32)
    Z, Y, X = v.shape
    own = ctx is None
    ctx = ctx or native.Context()
    try:
        labels = np.empty(X * Y * Z, np.int32)
        grad = np.empty(num_cells(X, Y, Z), np.uint8) if want_grad and stop is None else None
        m = None if mask is None else np.ascontiguousarray(mask, np.uint8).ravel()
        if m is not None and m.size != v.size:
            raise ValueError(f"mask holds {m.size} entries, the volume {v.size} vertices")
        st = native.LabelStats()
        flags = {None: 0, "regions": native.STOP_REGIONS, "integrated": native.STOP_INTEGRATED}[stop]
        flags |= native.FP32_FIRST if fp32_first else 0
        rc = native.lib().msceer_ondemand_asc(
            ctx.handle, v.ctypes.data_as(ctypes.c_void_p), X, Y, Z, err_thr, grad_thr, max_iters, presimp,
            None if m is None else m.ctypes.data_as(ctypes.c_void_p),
            None if grad is None else grad.ctypes.data_as(ctypes.c_void_p),
            labels.ctypes.data_as(ctypes.c_void_p), ctypes.byref(st), flags)
        native.check(rc, "msceer_ondemand_asc")
    finally:
        if own:
            ctx.close()
    return grad, labels.reshape(Z, Y, X), st.as_dict()


def ondemand(values: np.ndarray, need_asc1: bool = False, need_dsc1: bool = False, need_asc2: bool = True,
             err_thr: float = 1e-3, grad_thr: float = 1e-6, max_iters: int = 4096, presimp: float = 0.0,
             ctx: native.Context | None = None, need_dsc2: bool = False, with_dsc: bool = False):
    """OnDemandAccurate's accuracy branches (needASC1, needDSC1, needASC2,
    needDSC2): (conforming grad, labels (Z,Y,X) int32 or None without
    needASC2, restriction labels (cells int8), stats dict); with_dsc=True
    appends the descending labels (Z,Y,X) int32, or None without needDSC2."""
    v = np.ascontiguousarray(values, dtype=np.float32)
    Z, Y, X = v.shape
    need = ((native.NEED_ASC1 if need_asc1 else 0) | (native.NEED_DSC1 if need_dsc1 else 0) |
            (native.NEED_ASC2 if need_asc2 else 0) | (native.NEED_DSC2 if need_dsc2 else 0))
    own = ctx is None
    ctx = ctx or native.Context()
    try:
        labels = np.empty(X * Y * Z, np.int32) if need_asc2 else None
        labels_dsc = np.empty(X * Y * Z, np.int32) if need_dsc2 else None
        grad = np.empty(num_cells(X, Y, Z), np.uint8)
        restr = np.empty(num_cells(X, Y, Z), np.int8)
        st = native.LabelStats()
        rc = native.lib().msceer_ondemand2(
            ctx.handle, v.ctypes.data_as(ctypes.c_void_p), X, Y, Z, err_thr, grad_thr, max_iters, presimp, need,
            None, grad.ctypes.data_as(ctypes.c_void_p),
            None if labels is None else labels.ctypes.data_as(ctypes.c_void_p),
            None if labels_dsc is None else labels_dsc.ctypes.data_as(ctypes.c_void_p),
            restr.ctypes.data_as(ctypes.c_void_p), ctypes.byref(st), 0)
        native.check(rc, "msceer_ondemand2")
    finally:
        if own:
            ctx.close()
    out = (grad, None if labels is None else labels.reshape(Z, Y, X), restr, st.as_dict())
    if with_dsc:
        out += (None if labels_dsc is None else labels_dsc.reshape(Z, Y, X),)
    return out


def label_file_name(base: str, X: int, Y: int, Z: int) -> str:
    return f"{base}._int_{X}x{Y}x{Z}.raw"


def write_labels(path: str, labels: np.ndarray) -> None:
    np.ascontiguousarray(labels, dtype=np.int32).tofile(path)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def integrate_labels(values: np.ndarray, terminals: np.ndarray, zero_grad: np.ndarray | None = None,
                     err_thr: float = 1e-3, grad_thr: float = 1e-6, max_iters: int = 4096, sign: int = -1,
                     ctx: native.Context | None = None):
    """NumericIntegratorPathCompressingToTerminal::BeginIntegration on given terminals.

    terminals: int32 per vertex, -1 = integrate, else a certain terminal with that
    label; zero_grad: optional uint8 per vertex, gradient set to zero there;
    sign -1 flows downhill (ascending manifolds), +1 uphill.  Returns
    (labels (Z,Y,X) int32, stats dict).
    """
    v = np.ascontiguousarray(values, np.float32)
    Z, Y, X = v.shape
    t = np.ascontiguousarray(terminals, np.int32)
    if t.size != v.size:
        raise ValueError(f"terminals hold {t.size} entries, the volume {v.size} vertices")
    zg = None if zero_grad is None else np.ascontiguousarray(zero_grad, np.uint8)
    if zg is not None and zg.size != v.size:
        raise ValueError(f"zero_grad holds {zg.size} entries, the volume {v.size} vertices")
    out = np.empty(v.shape, np.int32)
    st = native.LabelStats()
    own = ctx is None
    ctx = ctx or native.Context()
    try:
        native.check(native.lib().msceer_integrate_labels(ctx.handle, _ptr(v), X, Y, Z, _ptr(t),
                                                          _ptr(zg) if zg is not None else None, float(err_thr),
                                                          float(grad_thr), int(max_iters), int(sign), _ptr(out),
                                                          ctypes.byref(st), 0), "msceer_integrate_labels")
    finally:
        if own:
            ctx.close()
    return out, st.as_dict()


def clean_labels(values: np.ndarray, labels: np.ndarray, ctx: native.Context | None = None) -> np.ndarray:
    """IsolatedCCRegionRemoverNEW::ComputeOutput on a copy of `labels`."""
    v = np.ascontiguousarray(values, np.float32)
    Z, Y, X = v.shape
    lab = np.ascontiguousarray(labels, np.int32).copy()
    if lab.size != v.size:
        raise ValueError(f"labels hold {lab.size} entries, the volume {v.size} vertices")
    own = ctx is None
    ctx = ctx or native.Context()
    try:
        native.check(native.lib().msceer_clean_labels(ctx.handle, _ptr(v), X, Y, Z, _ptr(lab), 0),
                     "msceer_clean_labels")
    finally:
        if own:
            ctx.close()
    return lab.reshape(Z, Y, X)


def restriction_labels(labels: np.ndarray, ctx: native.Context | None = None) -> np.ndarray:
    """ComputeMINBoundary: int8 per cell, shape (2Z-1, 2Y-1, 2X-1)."""
    lab = np.ascontiguousarray(labels, np.int32)
    Z, Y, X = lab.shape
    out = np.empty(num_cells(X, Y, Z), np.int8)
    own = ctx is None
    ctx = ctx or native.Context()
    try:
        native.check(native.lib().msceer_restriction_labels(ctx.handle, _ptr(lab), X, Y, Z, _ptr(out), 0),
                     "msceer_restriction_labels")
    finally:
        if own:
            ctx.close()
    return out.reshape(2 * Z - 1, 2 * Y - 1, 2 * X - 1)


def region_boundary_kind(labels: np.ndarray, ctx: native.Context | None = None) -> np.ndarray:
    """ComputeRegionBoundaryKind (Accurate3DMorseLabeling's `_edges_` file):
    int8 per cell, shape (2Z-1, 2Y-1, 2X-1)."""
    lab = np.ascontiguousarray(labels, np.int32)
    Z, Y, X = lab.shape
    out = np.empty(num_cells(X, Y, Z), np.int8)
    own = ctx is None
    ctx = ctx or native.Context()
    try:
        native.check(native.lib().msceer_region_boundary_kind(ctx.handle, _ptr(lab), X, Y, Z, _ptr(out), 0),
                     "msceer_region_boundary_kind")
    finally:
        if own:
            ctx.close()
    return out.reshape(2 * Z - 1, 2 * Y - 1, 2 * X - 1)


def restriction_add_max(values: np.ndarray, labels: np.ndarray, restriction: np.ndarray,
                        ctx: native.Context | None = None) -> np.ndarray:
    """ComputeMAXBoundary of `labels` added onto a copy of `restriction` (int8 per cell)."""
    v = np.ascontiguousarray(values, np.float32)
    Z, Y, X = v.shape
    lab = np.ascontiguousarray(labels, np.int32)
    if lab.size != v.size:
        raise ValueError(f"labels hold {lab.size} entries, the volume {v.size} vertices")
    r = np.ascontiguousarray(restriction, np.int8).ravel().copy()
    if r.size != num_cells(X, Y, Z):
        raise ValueError(f"restriction holds {r.size} entries, the cell complex {num_cells(X, Y, Z)} cells")
    own = ctx is None
    ctx = ctx or native.Context()
    try:
        native.check(native.lib().msceer_restriction_add_max(ctx.handle, _ptr(v), _ptr(lab), X, Y, Z, _ptr(r), 0),
                     "msceer_restriction_add_max")
    finally:
        if own:
            ctx.close()
    return r


def constrained_grad(values: np.ndarray, restriction: np.ndarray, ctx: native.Context | None = None,
                     skip_ldir: bool = False) -> np.ndarray:
    """Lower stars under a restriction labelling, then ldir and flags."""
    v = np.ascontiguousarray(values, np.float32)
    Z, Y, X = v.shape
    r = np.ascontiguousarray(restriction, np.int8)
    if r.size != num_cells(X, Y, Z):
        raise ValueError(f"restriction holds {r.size} entries, the cell complex {num_cells(X, Y, Z)} cells")
    out = np.empty(num_cells(X, Y, Z), np.uint8)
    own = ctx is None
    ctx = ctx or native.Context()
    try:
        native.check(native.lib().msceer_constrained_lstar(ctx.handle, _ptr(v), X, Y, Z, _ptr(r), _ptr(out),
                                                           native.SKIP_LDIR if skip_ldir else 0),
                     "msceer_constrained_lstar")
    finally:
        if own:
            ctx.close()
    return out.reshape(2 * Z - 1, 2 * Y - 1, 2 * X - 1)


def expanding_region_stop(values: np.ndarray, err: float = 1e-3, gthr: float = 1e-6, iters: int = 4096,
                          ctx: native.Context | None = None):
    """NumericIntegratorExpandingRegionStop (minima) of a (Z, Y, X) float32
    volume on the GPU: (labels int32 (Z, Y, X): vertex id of the minimum or
    -1, stats: minima, certain vertices, pass-1 streamlines, unambiguous
    vertices, re-integrated vertices)."""
    v = np.ascontiguousarray(values, np.float32)
    Z, Y, X = v.shape
    lab = np.empty(v.shape, np.int32)
    st = np.zeros(5, np.int64)
    own = ctx is None
    ctx = ctx or native.Context()
    try:
        native.check(native.lib().msceer_expanding_region_stop(ctx.handle, v.ctypes.data_as(ctypes.c_void_p), X, Y, Z,
                                                               err, gthr, iters,
                                                               lab.ctypes.data_as(ctypes.c_void_p),
                                                               st.ctypes.data_as(ctypes.c_void_p), 0),
                     "msceer_expanding_region_stop")
    finally:
        if own:
            ctx.close()
    return lab, st
