"""Batched volumes and brick-shape tuning on the GPU.

Every brick shape of the lower-star kernel gives the reference's bytes; a
batch gives each volume exactly what the single-volume call gives (the
.grad bit-exact; labels within the integrator's visiting-order tolerance)."""
import sys

import numpy as np
import pytest

from golden import load_lstar
from msceer_amd import batch, labels, lstar, native, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(0)
    yield c
    c.close()


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ["noise_48x97x31", "ties_10x11x12", "tiny_5x2x3", "smooth_60x50x40"])
@pytest.mark.parametrize("shape", [0, 1, 2, 3])
def test_every_brick_shape_is_bit_exact(ctx, name, shape):
    f, ref = load_lstar(name)
    native.check(native.lib().msceer_set_autotune(ctx.handle, 2 + shape), "msceer_set_autotune")
    try:
        g = lstar.steepest_lstar_grad(f, ctx=ctx).ravel()
    finally:
        native.lib().msceer_set_autotune(ctx.handle, 1)
    assert (g == ref).all(), int((g != ref).sum())


def test_autotune_picks_a_shape_and_stays_exact(ctx):
    f, ref = load_lstar("noise_48x97x31")
    native.lib().msceer_set_autotune(ctx.handle, 1)
    g = lstar.steepest_lstar_grad(f, ctx=ctx).ravel()
    assert (g == ref).all()
    assert native.lib().msceer_lstar_brick(ctx.handle, f.shape[2]) in (0, 1, 2, 3)


def _volumes(n, seed0=7000):
    out = []
    for i in range(n):
        s = synth.Stream(seed0 + i)
        u = s.uniform(3)
        shape = tuple(int(8 + 40 * float(x)) for x in u)
        out.append(synth.smoothed_noise(shape, seed0 + i, 1.5).astype(np.float32) if i % 2 == 0
                   else synth.gaussian_mixture(shape, seed0 + i, 6, 2.0, 6.0, -1.0, 1.0).astype(np.float32))
    return out


def test_batch_grads_equal_single_calls(ctx):
    vols = _volumes(12)
    grads, _, ms = batch.run_batch(vols, streams=4, ctx=ctx)
    for v, g in zip(vols, grads):
        assert (g == lstar.steepest_lstar_grad(v, ctx=ctx).ravel()).all()
    assert (ms > 0).all()


def test_batch_labels_equal_single_calls(ctx):
    vols = _volumes(8, 7100)
    grads, labs, _ = batch.run_batch(vols, labels=True, streams=3, ctx=ctx)
    for v, g, l in zip(vols, grads, labs):
        g1, l1, _ = labels.ondemand_asc(v, ctx=ctx)
        mism = float((l != l1).mean())
        _say(f"  {v.shape}: batch vs single label mismatch {mism:.5f}")
        # pass 1 labels path vertices by whichever streamline reaches them
        # first (as the reference's threads do), so runs differ slightly near
        # basin boundaries
        assert mism < 1e-2
        assert g.size == g1.size
