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
