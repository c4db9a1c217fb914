"""ctypes access to the CPU oracle (oracle/liboracle.so) -- the checker, never the product."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
_P = ctypes.c_void_p
_L = ctypes.c_longlong
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            import subprocess
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True)
        h = ctypes.CDLL(ORACLE_SO)
        h.oracle_maxmin.argtypes = [_P, _L, _L, _L, _P, _P]
        h.oracle_lstar_pairs.argtypes = [_P, _L, _L, _L, _P, _P]
        h.oracle_lstar_vertices.argtypes = [_P, _L, _L, _L, _P, _P, _L, _P]
        h.oracle_ldir.argtypes = [_P, _L, _L, _L]
        h.oracle_flags.argtypes = [_P, _L, _L, _L]
        h.oracle_steepest_grad.argtypes = [_P, _L, _L, _L, _P]
        h.oracle_set_threads.argtypes = [ctypes.c_int]
        _lib = h
    return _lib


def _p(a):
    return a.ctypes.data_as(_P)


def ncells(X, Y, Z):
    return (2 * X - 1) * (2 * Y - 1) * (2 * Z - 1)


def maxmin(f):
    f = np.ascontiguousarray(f, np.float32)
    Z, Y, X = f.shape
    mx = np.empty(ncells(X, Y, Z), np.uint8)
    mn = np.empty_like(mx)
    assert lib().oracle_maxmin(_p(f), X, Y, Z, _p(mx), _p(mn)) == 0
    return mx, mn


def lstar_pairs(f, restriction=None):
    f = np.ascontiguousarray(f, np.float32)
    Z, Y, X = f.shape
    g = np.empty(ncells(X, Y, Z), np.uint8)
    r = None if restriction is None else _p(np.ascontiguousarray(restriction, np.int8))
    assert lib().oracle_lstar_pairs(_p(f), X, Y, Z, r, _p(g)) == 0
    return g


def ldir(grad, shape_zyx):
    g = np.ascontiguousarray(grad, np.uint8).copy()
    Z, Y, X = shape_zyx
    assert lib().oracle_ldir(_p(g), X, Y, Z) == 0
    return g


def flags(grad, shape_zyx):
    g = np.ascontiguousarray(grad, np.uint8).copy()
    Z, Y, X = shape_zyx
    assert lib().oracle_flags(_p(g), X, Y, Z) == 0
    return g


def steepest_grad(f):
    f = np.ascontiguousarray(f, np.float32)
    Z, Y, X = f.shape
    g = np.empty(ncells(X, Y, Z), np.uint8)
    assert lib().oracle_steepest_grad(_p(f), X, Y, Z, _p(g)) == 0
    return g


def set_threads(n):
    return lib().oracle_set_threads(n)


# ---- accurate labelling (ondemandaccurate, ascending) ----------------------
_F = ctypes.c_float
_I = ctypes.c_int


def _integ_sigs():
    h = lib()
    if getattr(h, "_integ_ready", False):
        return h
    h.oracle_gradient_field.argtypes = [_P, _L, _L, _L, _P]
    h.oracle_certain_regions.argtypes = [_P, _L, _L, _L, _P, _F, _P, _P]
    h.oracle_certain_regions.restype = ctypes.c_int
    h.oracle_integrate.argtypes = [_P, _L, _L, _L, _P, _F, _F, _I, _I, _P, _P]
    h.oracle_integrate_signed.argtypes = [_P, _L, _L, _L, _P, _F, _F, _I, _I, _I, _P, _P]
    h.oracle_clean.argtypes = [_P, _L, _L, _L, _P]
    h.oracle_min_boundary.argtypes = [_P, _L, _L, _L, _P]
    h.oracle_max_boundary.argtypes = [_P, _P, _L, _L, _L, _P]
    h.oracle_conforming_vertices.argtypes = [_P, _L, _L, _L, _P, _P]
    h.oracle_conforming_vertices.restype = ctypes.c_longlong
    h.oracle_ondemand_asc.argtypes = [_P, _L, _L, _L, _F, _F, _I, _F, _I, _P, _P, _P, _P]
    h.oracle_region_boundary_kind.argtypes = [_P, _L, _L, _L, _P]
    h.oracle_saddle_lines.argtypes = [_P, _L, _L, _L, _P, _P, _F, _F, _I, _I, _I, _P, _P]
    h.oracle_ondemand.argtypes = [_P, _L, _L, _L, _F, _F, _I, _F, _I, _I, _I, _I, _P, _P, _P, _P]
    h._integ_ready = True
    return h


def gradient_field(f):
    f = np.ascontiguousarray(f, np.float32)
    Z, Y, X = f.shape
    g = np.empty((Z, Y, X, 3), np.float64)
    _integ_sigs().oracle_gradient_field(_p(f), X, Y, Z, _p(g))
    return g


def certain_regions(f, base_grad, presimp=0.0):
    f = np.ascontiguousarray(f, np.float32)
    Z, Y, X = f.shape
    lab = np.empty(X * Y * Z, np.int32)
    seeds = np.empty(X * Y * Z, np.int8)
    n = _integ_sigs().oracle_certain_regions(_p(f), X, Y, Z, _p(np.ascontiguousarray(base_grad, np.uint8)),
                                             presimp, _p(lab), _p(seeds))
    return lab, seeds, n


def integrate(f, certain, seeds, err=1e-3, gthr=1e-6, iters=4096, mode=0):
    """mode 0: the reference's mixed precision; mode 1: all FP64."""
    f = np.ascontiguousarray(f, np.float32)
    Z, Y, X = f.shape
    lab = np.ascontiguousarray(certain, np.int32).copy()
    st = np.zeros(8, np.int64)
    _integ_sigs().oracle_integrate(_p(f), X, Y, Z, _p(np.ascontiguousarray(seeds, np.int8)), err, gthr, iters,
                                   mode, _p(lab), _p(st))
    return lab, st


def integrate_signed(f, terminals, seeds, err=1e-3, gthr=1e-6, iters=4096, sign=-1, mode=0):
    """As integrate, flowing downhill (sign -1) or uphill (+1)."""
    f = np.ascontiguousarray(f, np.float32)
    Z, Y, X = f.shape
    lab = np.ascontiguousarray(terminals, np.int32).ravel().copy()
    st = np.zeros(8, np.int64)
    sd = None if seeds is None else np.ascontiguousarray(seeds, np.int8).ravel()
    rc = _integ_sigs().oracle_integrate_signed(_p(f), X, Y, Z, _p(sd) if sd is not None else None, err, gthr, iters,
                                               sign, mode, _p(lab), _p(st))
    assert rc == 0
    return lab, st


def clean(f, labels):
    f = np.ascontiguousarray(f, np.float32)
    Z, Y, X = f.shape
    lab = np.ascontiguousarray(labels, np.int32).copy()
    _integ_sigs().oracle_clean(_p(f), X, Y, Z, _p(lab))
    return lab


def min_boundary(labels, shape_zyx):
    Z, Y, X = shape_zyx
    r = np.empty(ncells(X, Y, Z), np.int8)
    _integ_sigs().oracle_min_boundary(_p(np.ascontiguousarray(labels, np.int32)), X, Y, Z, _p(r))
    return r


def max_boundary(f, labels, restriction):
    """ComputeMAXBoundary of `labels` added onto a copy of `restriction`."""
    f = np.ascontiguousarray(f, np.float32)
    Z, Y, X = f.shape
    r = np.ascontiguousarray(restriction, np.int8).ravel().copy()
    _integ_sigs().oracle_max_boundary(_p(f), _p(np.ascontiguousarray(labels, np.int32)), X, Y, Z, _p(r))
    return r


def conforming_vertices(f, restriction):
    f = np.ascontiguousarray(f, np.float32)
    Z, Y, X = f.shape
    r = np.ascontiguousarray(restriction, np.int8)
    h = _integ_sigs()
    n = h.oracle_conforming_vertices(_p(f), X, Y, Z, _p(r), None)
    out = np.empty(n, np.int64)
    h.oracle_conforming_vertices(_p(f), X, Y, Z, _p(r), _p(out))
    return out


def ondemand_asc(f, err=1e-3, gthr=1e-6, iters=4096, presimp=0.0, mode=0):
    """(grad, cleaned labels, restriction, stats[traced, over_budget, refixed, conforming])."""
    f = np.ascontiguousarray(f, np.float32)
    Z, Y, X = f.shape
    g = np.empty(ncells(X, Y, Z), np.uint8)
    lab = np.empty(X * Y * Z, np.int32)
    r = np.empty(ncells(X, Y, Z), np.int8)
    st = np.zeros(8, np.int64)
    _integ_sigs().oracle_ondemand_asc(_p(f), X, Y, Z, err, gthr, iters, presimp, mode, _p(g), _p(lab), _p(r), _p(st))
    return g, lab, r, st


def ondemand(f, need_asc1, need_dsc1, need_asc2, err=1e-3, gthr=1e-6, iters=4096, presimp=0.0, mode=0):
    """OnDemandAccurate with the needASC1 / needDSC1 / needASC2 flags:
    (grad, labels (all -1 unless need_asc2), restriction, stats[20]: [0..3]
    labelling, [4..11] ASC lines, [12..19] DSC lines)."""
    f = np.ascontiguousarray(f, np.float32)
    Z, Y, X = f.shape
    g = np.empty(ncells(X, Y, Z), np.uint8)
    lab = np.full(X * Y * Z, -1, np.int32)
    r = np.empty(ncells(X, Y, Z), np.int8)
    st = np.zeros(20, np.int64)
    _integ_sigs().oracle_ondemand(_p(f), X, Y, Z, err, gthr, iters, presimp, int(need_asc1), int(need_dsc1),
                                  int(need_asc2), mode, _p(g), _p(lab), _p(r), _p(st))
    return g, lab, r, st


def region_boundary_kind(labels, shape_zyx):
    """ComputeRegionBoundaryKind of X*Y*Z int32 labels: int8 per cell."""
    Z, Y, X = shape_zyx
    lab = np.ascontiguousarray(labels, np.int32).ravel()
    out = np.empty(ncells(X, Y, Z), np.int8)
    _integ_sigs().oracle_region_boundary_kind(_p(lab), X, Y, Z, _p(out))
    return out


def saddle_lines(f, grad, asc, restriction=None, seeds=None, err=1e-3, gthr=1e-6, iters=4096, mode=0):
    """The digitized 1-manifolds of one family added onto restriction
    (a copy): (restriction, stats[8]: saddles, streamlines, voxel traversals,
    cells labelled, stops on labels or the boundary, low-gradient stops,
    budget stops, microseconds of the saddle loop)."""
    f = np.ascontiguousarray(f, np.float32)
    Z, Y, X = f.shape
    r = np.zeros(ncells(X, Y, Z), np.int8) if restriction is None else np.array(restriction, np.int8)
    sd = None if seeds is None else np.ascontiguousarray(seeds, np.int8)
    st = np.zeros(8, np.int64)
    _integ_sigs().oracle_saddle_lines(_p(f), X, Y, Z, _p(np.ascontiguousarray(grad, np.uint8)),
                                      _p(sd) if sd is not None else None, err, gthr, iters, int(asc), mode, _p(r),
                                      _p(st))
    return r, st


# ---- near-separatrix classifier ----------------------------------------------
def _cnn_sigs():
    h = lib()
    if getattr(h, "_cnn_ready", False):
        return h
    h.oracle_cnn_num_params.restype = ctypes.c_int
    h.oracle_cnn_normalize.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong,
                                       ctypes.c_void_p]
    h.oracle_cnn_predict.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_longlong,
                                     ctypes.c_longlong, ctypes.c_int, ctypes.c_void_p]
    h.oracle_cnn_gradient.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_longlong,
                                      ctypes.c_longlong, ctypes.c_void_p, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_void_p, ctypes.c_void_p]
    h._cnn_ready = True
    return h


def cnn_num_params():
    return _cnn_sigs().oracle_cnn_num_params()


def cnn_normalize(f):
    f = np.ascontiguousarray(f, np.float32)
    Z, Y, X = f.shape
    out = np.empty_like(f)
    _cnn_sigs().oracle_cnn_normalize(_p(f), X, Y, Z, _p(out))
    return out


def cnn_predict(params, f, quant=1):
    f = np.ascontiguousarray(f, np.float32)
    p = np.ascontiguousarray(params, np.float32)
    Z, Y, X = f.shape
    out = np.empty(f.shape, np.float64)
    _cnn_sigs().oracle_cnn_predict(_p(p), _p(f), X, Y, Z, quant, _p(out))
    return out


def cnn_gradient(params, f, target, pos_w=1.0, quant=1):
    """(mean weighted BCE, gradient (float64))."""
    f = np.ascontiguousarray(f, np.float32)
    p = np.ascontiguousarray(params, np.float32)
    t = np.ascontiguousarray(target, np.uint8)
    Z, Y, X = f.shape
    loss = np.zeros(1, np.float64)
    grad = np.zeros(cnn_num_params(), np.float64)
    _cnn_sigs().oracle_cnn_gradient(_p(p), _p(f), X, Y, Z, _p(t), float(pos_w), quant, _p(loss), _p(grad))
    return float(loss[0]), grad


def msc_base(grad, shape_zyx):
    """extractmsc's base reading of a .grad: (basins, mounts, counts) -- int32
    node ids per vertex (nodes = critical cells in cell-id order) and the
    critical cells by dimension."""
    h = lib()
    h.oracle_msc_base.argtypes = [_P, _L, _L, _L, _P, _P, _P]
    h.oracle_msc_base.restype = ctypes.c_longlong
    g = np.ascontiguousarray(grad, np.uint8).ravel()
    Z, Y, X = shape_zyx
    basins = np.empty(X * Y * Z, np.int32)
    mounts = np.empty(X * Y * Z, np.int32)
    counts = np.zeros(4, np.int64)
    h.oracle_msc_base(_p(g), X, Y, Z, _p(basins), _p(mounts), _p(counts))
    return basins.reshape(shape_zyx), mounts.reshape(shape_zyx), counts


def expanding_region_stop(f, err, gthr, iters, mode=0):
    """NumericIntegratorExpandingRegionStop (minima): (labels int32 (Z,Y,X),
    stats[5]).  mode 0: reference arithmetic and one thread's order; 1: the
    same in FP64; 2: FP64, every vertex traced alone, the re-integration
    reading pass 1's labels."""
    h = lib()
    h.oracle_expanding_region_stop.argtypes = [_P, _L, _L, _L, _F, _F, _I, _I, _P, _P]
    f = np.ascontiguousarray(f, np.float32)
    Z, Y, X = f.shape
    lab = np.empty(f.shape, np.int32)
    st = np.zeros(5, np.int64)
    h.oracle_expanding_region_stop(_p(f), X, Y, Z, float(err), float(gthr), int(iters), int(mode), _p(lab), _p(st))
    return lab, st


def ondemand2(f, err, gthr, iters, presimp, asc1=0, dsc1=0, asc2=1, dsc2=0, mode=0):
    """OnDemandAccurate with any accuracy branches: (grad, labels_asc,
    labels_dsc, restriction, stats[20])."""
    h = lib()
    h.oracle_ondemand2.argtypes = [_P, _L, _L, _L, _F, _F, _I, _F, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P]
    f = np.ascontiguousarray(f, np.float32)
    Z, Y, X = f.shape
    grad = np.empty(ncells(X, Y, Z), np.uint8)
    la = np.full(f.size, -1, np.int32)
    ld = np.full(f.size, -1, np.int32)
    r = np.empty(ncells(X, Y, Z), np.int8)
    st = np.zeros(20, np.int64)
    h.oracle_ondemand2(_p(f), X, Y, Z, float(err), float(gthr), int(iters), float(presimp), int(asc1), int(dsc1),
                       int(asc2), int(dsc2), int(mode), _p(grad), _p(la), _p(ld), _p(r), _p(st))
    return grad, la.reshape(f.shape), ld.reshape(f.shape), r, st
