"""Batched volumes and brick-shape tuning on the GPU.

Every brick shape of the lower-star kernel gives the reference's bytes; a
batch gives each volume exactly what the single-volume call gives (the
labels do not depend on scheduling, so they are identical too).  The
512-volume workload's family (sides 32..160, smoothed noise and Gaussian
mixtures) is checked on 16 of its volumes that span its sizes: the conforming
.grad of the batched labels bit-exact against the oracle's, and the labels
within 0.1 % of every vertex's exact FP64 destination (oracle mode 2), and no
further from it than the reference's own algorithm in its own arithmetic
(oracle mode 0, path compression in one thread's order) plus 0.05 %."""
import sys

import numpy as np
import pytest

import oracle_lib
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
        _say(f"  {v.shape}: batch vs single label differences {int((l != l1).sum())}")
        assert (l == l1).all()
        assert (g == g1.ravel()).all()


def _c4_subset():
    """16 volumes of the batched workload spanning its sizes (every 32nd by
    vertex count, smallest first)."""
    sizes = []
    for i in range(512):
        u = synth.Stream(4000 + i).uniform(5)
        shape = tuple(int(32 + min(128, int(129 * float(u[k])))) for k in range(3))
        sizes.append((shape[0] * shape[1] * shape[2], i))
    sizes.sort()
    return [i for _, i in sizes[::32]]


C4_SUBSET = _c4_subset()


@pytest.fixture(scope="module")
def c4_batch(ctx):
    vols = [synth.batch_volume(i).astype(np.float32) for i in C4_SUBSET]
    _say(f"c4 batch: {len(vols)} volumes, {sum(v.size for v in vols)} vertices, shapes "
         f"{[v.shape for v in vols]}")
    grads, labs, ms = batch.run_batch(vols, labels=True, streams=4, ctx=ctx)
    _say(f"  per-volume ms {np.round(ms, 1).tolist()}")
    return vols, grads, labs


@pytest.mark.parametrize("k", range(len(C4_SUBSET)))
def test_c4_batch_volume_against_oracle(c4_batch, k):
    vols, grads, labs = c4_batch
    v, g, lab = vols[k], grads[k], labs[k]
    _say(f"c4 volume {C4_SUBSET[k]} {v.shape}")
    # the conforming .grad of the batched labels, as the oracle builds it
    r = oracle_lib.min_boundary(lab.ravel(), v.shape)
    expect = oracle_lib.flags(oracle_lib.ldir(oracle_lib.lstar_pairs(v, r), v.shape), v.shape)
    assert (g == expect).all(), int((g != expect).sum())
    _, exact, _, _ = oracle_lib.ondemand_asc(v, 1e-3, 1e-6, 4096, 0.0, mode=2)
    _, lab_ref, _, _ = oracle_lib.ondemand_asc(v, 1e-3, 1e-6, 4096, 0.0, mode=0)
    mism = float((lab.ravel() != exact).mean())
    ref_mism = float((lab_ref != exact).mean())
    _say(f"  label mismatch vs exact FP64 destinations {mism:.6f} (reference algorithm {ref_mism:.6f})")
    assert mism <= 1e-3
    assert mism <= ref_mism + 5e-4


def test_c4_batch_lower_star_bit_exact(ctx):
    vols = [synth.batch_volume(i).astype(np.float32) for i in C4_SUBSET[::3]]
    grads, _, _ = batch.run_batch(vols, streams=4, ctx=ctx)
    for v, g in zip(vols, grads):
        ref = oracle_lib.steepest_grad(v)
        assert (g == ref).all(), (v.shape, int((g != ref).sum()))
