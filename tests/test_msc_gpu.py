"""The GPU reading of a .grad as extractmsc reads it (msceer_msc_base: nodes,
base basins, base mountains), bit-exact against the reference's own
Morse-Smale complex class on the lower-star goldens and against the oracle
on GPU-made .grad files of larger seeded volumes."""
import ctypes
import sys

import numpy as np
import pytest

import oracle_lib
from golden import load_msc, msc_case_names
from msceer_amd import hipmem, lstar, msc, native, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name", msc_case_names())
def test_reference_goldens(ctx, name):
    f, grad, basins, mounts, counts = load_msc(name)
    b, m, c = msc.msc_base(grad, f.shape, ctx=ctx)
    assert (c == counts).all(), (c, counts)
    assert (b == basins).all(), int((b != basins).sum())
    assert (m == mounts).all(), int((m != mounts).sum())


@pytest.mark.parametrize("shape,seed,kind", [((96, 112, 128), 41, "noise"), ((128, 128, 128), 42, "gauss"),
                                             ((7, 300, 33), 43, "noise")])
def test_gpu_grad_against_oracle(ctx, shape, seed, kind):
    f = (synth.smoothed_noise(shape, seed, 2.0) if kind == "noise"
         else synth.gaussian_mixture(shape, seed, 40, 3.0, 10.0, -1.0, 1.0)).astype(np.float32)
    g = lstar.steepest_lstar_grad(f, ctx=ctx).ravel()
    print(f"{shape} {kind}", flush=True)
    sys.stdout.flush()
    b, m, c = msc.msc_base(g, f.shape, ctx=ctx)
    rb, rm, rc = oracle_lib.msc_base(g, f.shape)
    assert (c == rc).all()
    assert (b == rb).all(), int((b != rb).sum())
    assert (m == rm).all(), int((m != rm).sum())


def test_device_pointers_and_one_output(ctx):
    f, grad, basins, mounts, counts = load_msc("noise_48x97x31")
    Z, Y, X = f.shape
    dg = hipmem.DeviceBuffer(grad.size)
    dg.upload(grad)
    dm = hipmem.DeviceBuffer(4 * f.size)
    dm.fill(0x7F)
    hipmem.synchronize()
    native.check(native.lib().msceer_msc_base(ctx.handle, ctypes.c_void_p(dg.addr), X, Y, Z, None,
                                              ctypes.c_void_p(dm.addr), None, native.DEVICE_PTRS), "msceer_msc_base")
    hipmem.synchronize()
    m = dm.download(np.int32, f.size).reshape(f.shape)
    assert (m == mounts).all()
