"""msceer_lstar_grad_noalloc (the .grad with every scratch buffer in a static
device array) against msceer_lstar_grad on the same device buffers: two calls
in a row, the second on a smaller volume, bit-exact; volumes whose scratch
would not fit are refused."""
import ctypes
import sys

import numpy as np
import pytest

import oracle_lib
from msceer_amd import hipmem, native, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _grad(fn, ctx, f, flags=0):
    Z, Y, X = f.shape
    C = (2 * X - 1) * (2 * Y - 1) * (2 * Z - 1)
    dv = hipmem.DeviceBuffer(f.nbytes)
    dv.upload(f)
    dg = hipmem.DeviceBuffer(C)
    dg.fill(0)
    rc = fn(ctx.handle, ctypes.c_void_p(dv.addr), X, Y, Z, ctypes.c_void_p(dg.addr), flags)
    hipmem.synchronize()
    return rc, dg.download(np.uint8, C)


def test_noalloc_matches_the_allocating_entry_point():
    lib = native.lib()
    ctx = native.Context(0)
    try:
        for shape, seed in (((64, 96, 112), 31), ((23, 40, 37), 32)):
            f = synth.smoothed_noise(shape, seed, 1.5)
            _say(f"noalloc {shape}")
            rc, got = _grad(lib.msceer_lstar_grad_noalloc, ctx, f)
            native.check(rc, "msceer_lstar_grad_noalloc")
            rc, ref = _grad(lib.msceer_lstar_grad, ctx, f, native.DEVICE_PTRS)
            native.check(rc, "msceer_lstar_grad")
            assert (got == ref).all(), int((got != ref).sum())
            assert (got == oracle_lib.steepest_grad(f)).all()
            rc, nold = _grad(lib.msceer_lstar_grad_noalloc, ctx, f, native.SKIP_LDIR)
            native.check(rc, "msceer_lstar_grad_noalloc")
            _, ref2 = _grad(lib.msceer_lstar_grad, ctx, f, native.DEVICE_PTRS | native.SKIP_LDIR)
            assert (nold == ref2).all()
    finally:
        ctx.close()


def test_noalloc_refuses_what_does_not_fit():
    lib = native.lib()
    ctx = native.Context(0)
    try:
        # 128^3 vertices need more scratch than the static array holds; the
        # call returns before touching the (null) buffers
        assert lib.msceer_lstar_grad_noalloc(ctx.handle, ctypes.c_void_p(0), 128, 128, 128, ctypes.c_void_p(0),
                                             0) == native.EINVAL
        rc = lib.msceer_lstar_grad_noalloc(ctx.handle, ctypes.c_void_p(1), 128, 128, 128, ctypes.c_void_p(1), 0)
        assert rc == native.ETOOBIG
        assert lib.msceer_lstar_grad_noalloc(ctx.handle, ctypes.c_void_p(1), 8, 8, 8, ctypes.c_void_p(1),
                                             native.DEVICE_PTRS) == native.EINVAL
    finally:
        ctx.close()
