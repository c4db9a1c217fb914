"""ctypes access to tests/emul/libemul.so: the lower-star kernel's per-lane
logic (msceer_amd/csrc/lstar_core.hpp) built for the host."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "emul", "libemul.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        subprocess.run(["make", "-s", "-C", os.path.join(HERE, "emul")], check=True)
        h = ctypes.CDLL(SO)
        P, L = ctypes.c_void_p, ctypes.c_longlong
        h.emul_lstar_prepass.argtypes = [P, L, L, L, P, P]
        h.emul_check_product.argtypes = [P, L, L, L, P]
        _lib = h
    return _lib


def lstar_prepass(f, restriction=None):
    f = np.ascontiguousarray(f, np.float32)
    Z, Y, X = f.shape
    g = np.zeros((2 * X - 1) * (2 * Y - 1) * (2 * Z - 1), np.uint8)
    r = None if restriction is None else np.ascontiguousarray(restriction, np.int8).ctypes.data_as(ctypes.c_void_p)
    lib().emul_lstar_prepass(f.ctypes.data_as(ctypes.c_void_p), X, Y, Z, r, g.ctypes.data_as(ctypes.c_void_p))
    return g


def expected_prepass(pairs_with_flags, shape_zyx):
    """Oracle pairs+flags with the kernel's ldir preset (3; 0 for critical 3-cells)."""
    import gradcheck
    Z, Y, X = shape_zyx
    dims, _ = gradcheck.cell_dims(X, Y, Z)
    crit3 = (((pairs_with_flags >> 2) & 7) == 7) & (dims == 3)
    return pairs_with_flags | np.where(crit3, 0, 3 << 5).astype(np.uint8)


def check_product(f):
    """(interior vertices whose product lower star the closed form expanded,
    of those the number whose bytes differ from the general expansion's)."""
    f = np.ascontiguousarray(f, np.float32)
    Z, Y, X = f.shape
    c = np.zeros(2, np.int64)
    lib().emul_check_product(f.ctypes.data_as(ctypes.c_void_p), X, Y, Z, c.ctypes.data_as(ctypes.c_void_p))
    return int(c[0]), int(c[1])


# ---- accurate 1-manifolds (msceer_amd/csrc/lines_core.hpp) ------------------
LINES_SO = os.path.join(HERE, "emul", "liblines_emul.so")
_lines = None


def lines_lib():
    global _lines
    if _lines is None:
        subprocess.run(["make", "-s", "-C", os.path.join(HERE, "emul")], check=True)
        h = ctypes.CDLL(LINES_SO)
        P, L = ctypes.c_void_p, ctypes.c_longlong
        h.emul_lines.argtypes = [P, L, L, L, P, P, P, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                 L, ctypes.c_int, P]
        _lines = h
    return _lines


def lines(f, grad, field, asc, window=1 << 20, order=0, err=1e-3, gthr=1e-6, iters=4096):
    """The GPU's window algorithm for one 1-manifold family, run on the host:
    (restriction labels, stats[8]: saddles, windows, cuts, lines traced alone,
    speculative traces, replayed traces, replays stopped on a foreign test /
    a foreign label).  field: the FP64 gradient (Z, Y, X, 3); order: the
    order in which a launch's lanes run (0 ascending, 1 descending, 2 random)."""
    f = np.ascontiguousarray(f, np.float32)
    Z, Y, X = f.shape
    fld = np.ascontiguousarray(field, np.float64)
    g = np.ascontiguousarray(grad, np.uint8)
    r = np.zeros((2 * X - 1) * (2 * Y - 1) * (2 * Z - 1), np.int8)
    st = np.zeros(8, np.int64)
    rc = lines_lib().emul_lines(f.ctypes.data_as(ctypes.c_void_p), X, Y, Z, g.ctypes.data_as(ctypes.c_void_p),
                                fld.ctypes.data_as(ctypes.c_void_p), r.ctypes.data_as(ctypes.c_void_p), int(asc),
                                err, gthr, iters, window, order, st.ctypes.data_as(ctypes.c_void_p))
    if rc:
        raise RuntimeError(f"emul_lines returned {rc}")
    return r, st
