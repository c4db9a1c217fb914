"""The .grad file that the tiled program writes (save_discrete_gradient_to_file1,
steepest_lstar_tiled/localized_morse_complex.h:1063-1133): the tiles' rows in
global row-major order, each byte holding the pairing (assigned bit, pair
direction) and the boundary flag set on the domain's boundary cells, the
ascending-manifold dimension left unset (a TODO there).  That is the
steepest_lstar .grad with its ldir bits (5-7) cleared, which
msceer_lstar_grad writes with MSCEER_TILED_LAYOUT.  The CPU test pins the
layout on the oracle against the reference goldens; the GPU test checks the
flag, on host and on odd-offset device buffers (the clearing pass takes
aligned words and the bytes either side)."""
import numpy as np
import pytest

import oracle_lib
from golden import lstar_case_names, load_lstar
from msceer_amd import lstar, native


@pytest.mark.parametrize("name", lstar_case_names())
def test_pairs_and_flags_are_the_golden_without_ldir(name):
    f, ref = load_lstar(name)
    g = oracle_lib.flags(oracle_lib.lstar_pairs(f), f.shape)
    assert (g.ravel() == (ref.ravel() & 0x1F)).all()


@pytest.mark.gpu
@pytest.mark.parametrize("name", lstar_case_names())
def test_tiled_layout_flag(name):
    import ctypes
    from msceer_amd import hipmem
    f, ref = load_lstar(name)
    expect = ref.ravel() & 0x1F
    ctx = native.Context(0)
    try:
        g = lstar.steepest_lstar_grad(f, ctx=ctx, tiled_layout=True).ravel()
        assert (g == expect).all(), int((g != expect).sum())
        # device buffers at an odd offset
        Z, Y, X = f.shape
        dv = hipmem.DeviceBuffer(f.nbytes)
        dv.upload(np.ascontiguousarray(f, np.float32))
        dg = hipmem.DeviceBuffer(expect.size + 8)
        dg.fill(0xFF)
        rc = native.lib().msceer_lstar_grad(ctx.handle, ctypes.c_void_p(dv.addr), X, Y, Z,
                                            ctypes.c_void_p(dg.addr + 3), native.DEVICE_PTRS | native.TILED_LAYOUT)
        native.check(rc, "msceer_lstar_grad")
        hipmem.synchronize()
        raw = dg.download(np.uint8, expect.size + 8)
        assert (raw[3:3 + expect.size] == expect).all()
        assert (raw[:3] == 0xFF).all() and (raw[3 + expect.size:] == 0xFF).all()
    finally:
        ctx.close()
