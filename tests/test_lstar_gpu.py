"""GPU parity of the lower-star gradient (through the C ABI) against the
reference golden vectors and the CPU oracle.  Bar: bit-exact .grad bytes."""
import sys

import numpy as np
import pytest

import oracle_lib
from golden import lstar_case_names, load_lstar
from msceer_amd import hipmem, lstar, native, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(0)
    yield c
    c.close()


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", lstar_case_names())
def test_golden_bit_exact(ctx, name):
    f, ref = load_lstar(name)
    _say(f"lstar golden {name} {f.shape}")
    g = lstar.steepest_lstar_grad(f, ctx).ravel()
    bad = np.nonzero(g != ref)[0]
    assert bad.size == 0, f"{name}: {bad.size} bytes differ, first cells {bad[:8]}"


@pytest.mark.parametrize("shape,seed,kind", [
    ((33, 17, 70), 1, "normal"),
    ((20, 41, 129), 2, "smooth"),
    ((64, 3, 65), 3, "ties"),
    ((7, 130, 66), 4, "normal"),
])
def test_oracle_parity_seeded(ctx, shape, seed, kind):
    _say(f"lstar seeded {shape} {seed} {kind}")
    s = synth.Stream(seed)
    n = int(np.prod(shape))
    if kind == "normal":
        f = s.normal(n).reshape(shape).astype(np.float32)
    elif kind == "ties":
        f = np.floor(s.uniform(n) * 4).reshape(shape).astype(np.float32)
    else:
        f = synth.smoothed_noise(shape, seed, 1.5)
    ref = oracle_lib.steepest_grad(f)
    g = lstar.steepest_lstar_grad(f, ctx).ravel()
    assert (g == ref).all(), int((g != ref).sum())


def test_skip_ldir_matches_pairs_with_preset_ldir(ctx):
    f, _ = load_lstar("smooth_60x50x40")
    pairs = oracle_lib.flags(oracle_lib.lstar_pairs(f), f.shape)
    crit3 = (((pairs >> 2) & 7) == 7)
    Z, Y, X = f.shape
    import gradcheck
    dims, _ = gradcheck.cell_dims(X, Y, Z)
    expect = pairs | np.where(crit3 & (dims == 3), 0, 3 << 5).astype(np.uint8)
    g = lstar.steepest_lstar_grad(f, ctx, skip_ldir=True).ravel()
    assert (g == expect).all()


def test_device_pointers_write_every_cell_once(ctx):
    f, ref = load_lstar("c1_gauss64")
    Z, Y, X = f.shape
    dv = hipmem.DeviceBuffer(f.nbytes)
    dv.upload(f)
    dg = hipmem.DeviceBuffer(ref.size)
    dg.fill(0xFF)  # 0xFF has ldir 7, which never occurs: any unwritten byte shows
    lstar.steepest_lstar_grad_device(ctx, dv.addr, X, Y, Z, dg.addr)
    g = dg.download(np.uint8, ref.size)
    assert (g == ref).all()


def test_repeat_calls_identical_and_timing(ctx):
    f, ref = load_lstar("smooth_96")
    a = lstar.steepest_lstar_grad(f, ctx, timing=True).ravel()
    t = ctx.last_timing()
    b = lstar.steepest_lstar_grad(f, ctx).ravel()
    assert (a == ref).all() and (b == ref).all()
    assert t.lstar_ms > 0 and t.total_ms >= t.lstar_ms
    crit = ((ref >> 2) & 7) == 7
    import gradcheck
    dims, _ = gradcheck.cell_dims(96, 96, 96)
    assert [int((crit & (dims == k)).sum()) for k in range(4)] == list(t.critical)


def test_full_size_512_matches_oracle(ctx):
    """512^3 smoothed noise (sigma 2, seed 202): the benchmark volume."""
    _say("lstar 512^3 full size")
    f = synth.lstar_large(202, 512)
    g = lstar.steepest_lstar_grad(f, ctx, timing=True).ravel()
    t = ctx.last_timing()
    _say(f"gpu lstar {t.lstar_ms:.3f} ms ldir {t.ldir_ms:.3f} ms rounds {t.ldir_rounds}")
    crit = list(t.critical)
    assert crit[0] - crit[1] + crit[2] - crit[3] == 1
    oracle_lib.set_threads(16)
    ref = oracle_lib.steepest_grad(f)
    bad = int((g != ref).sum())
    assert bad == 0, bad
