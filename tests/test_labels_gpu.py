"""GPU accurate labelling (msceer_ondemand_asc) against the reference goldens
and the CPU oracle.

Stages that follow the reference operation for operation are bit-exact:
certain regions, and the conforming .grad given the labels.  The integrated
labels are compared as a mismatch rate with the FP64 oracle, at most 0.1 % of
the vertices (pass 1 of the reference depends on its threads' visiting
order, the GPU's levels are synchronous, and the GPU traces in FP32 first).
The GPU labels themselves are the same from run to run.
"""
import sys

import numpy as np
import pytest

import gradcheck
import oracle_lib
from golden import load_ondemand, ondemand_case_names
from msceer_amd import labels, native

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(0)
    yield c
    c.close()


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ondemand_case_names())
def test_certain_regions_bit_exact(ctx, name):
    f, (err, gthr, iters, presimp), ref = load_ondemand(name)
    _say(f"labels regions {name} {f.shape}")
    _, lab, st = labels.ondemand_asc(f, err, gthr, iters, presimp, ctx=ctx, stop="regions")
    assert (lab.ravel() == ref["certain"]).all(), int((lab.ravel() != ref["certain"]).sum())


@pytest.mark.parametrize("name", ondemand_case_names())
def test_labels_and_conforming_grad(ctx, name):
    f, (err, gthr, iters, presimp), ref = load_ondemand(name)
    Z, Y, X = f.shape
    _say(f"labels full {name} {f.shape}")
    grad, lab, st = labels.ondemand_asc(f, err, gthr, iters, presimp, ctx=ctx)
    lab = lab.ravel()
    _, lab64, _, _ = oracle_lib.ondemand_asc(f, err, gthr, iters, presimp, mode=1)
    mism = float((lab != lab64).mean())
    _say(f"  mismatch vs FP64 oracle {mism:.5f}, vs reference {float((lab != ref['cleaned']).mean()):.5f}, {st}")
    assert mism <= 1e-3
    # the .grad is a pure function of the labels: bit-exact
    r = oracle_lib.min_boundary(lab, f.shape)
    expect = oracle_lib.flags(oracle_lib.ldir(oracle_lib.lstar_pairs(f, r), f.shape), f.shape)
    assert (grad == expect).all(), int((grad != expect).sum())
    gradcheck.check_consistency(grad, X, Y, Z)


@pytest.mark.parametrize("shape", [(2, 2, 2), (5, 2, 3), (2, 7, 5), (4, 3, 9)])
def test_tiny_volumes(ctx, shape):
    # the smallest grids (two vertices along an axis, one tile, one stride):
    # the pipeline runs, labels are minima of the field's own gradient or -1,
    # and the conforming .grad is the oracle's for those labels
    rng = np.random.default_rng(sum(shape) * 7)
    f = rng.standard_normal(shape).astype(np.float32)
    Z, Y, X = shape
    _say(f"labels tiny {shape}")
    grad, lab, _ = labels.ondemand_asc(f, ctx=ctx)
    lab = lab.ravel()
    assert ((lab == -1) | ((lab >= 0) & (lab < f.size))).all()
    r = oracle_lib.min_boundary(lab, f.shape)
    expect = oracle_lib.flags(oracle_lib.ldir(oracle_lib.lstar_pairs(f, r), f.shape), f.shape)
    assert (grad == expect).all(), int((grad != expect).sum())
    gradcheck.check_consistency(grad, X, Y, Z)


@pytest.mark.parametrize("kind", ["constant", "ramp"])
def test_degenerate_fields(ctx, kind):
    # all ties (order by vertex id) and a linear ramp: the box's corners are
    # critical in their own boundary strata; the labels equal the FP64
    # oracle's exactly (every vertex is certain or flows straight down)
    Z, Y, X = 11, 13, 17
    zz, yy, xx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    f = np.zeros((Z, Y, X), np.float32) if kind == "constant" else (xx + 2.0 * yy + 3.0 * zz).astype(np.float32)
    _say(f"labels {kind} {f.shape}")
    grad, lab, st = labels.ondemand_asc(f, ctx=ctx)
    _, lab64, _, _ = oracle_lib.ondemand_asc(f, 1e-3, 1e-6, 4096, 0.0, mode=1)
    assert (lab.ravel() == lab64).all(), int((lab.ravel() != lab64).sum())
    r = oracle_lib.min_boundary(lab.ravel(), f.shape)
    expect = oracle_lib.flags(oracle_lib.ldir(oracle_lib.lstar_pairs(f, r), f.shape), f.shape)
    assert (grad == expect).all(), int((grad != expect).sum())


def test_stages_are_monotone(ctx):
    # certain regions keep their labels through integration and cleanup
    f, (err, gthr, iters, presimp), ref = load_ondemand("smooth_34x33x32")
    _, cert, _ = labels.ondemand_asc(f, err, gthr, iters, presimp, ctx=ctx, stop="regions")
    _, integ, _ = labels.ondemand_asc(f, err, gthr, iters, presimp, ctx=ctx, stop="integrated")
    c, i = cert.ravel(), integ.ravel()
    assert (i[c != -1] == c[c != -1]).all()
    assert float((i == -1).mean()) < 0.01


def test_mask_outside_takes_combinatorial_labels(ctx):
    f, (err, gthr, iters, presimp), ref = load_ondemand("labels_48")
    mask = np.zeros(f.shape, np.uint8)
    _, lab, st = labels.ondemand_asc(f, err, gthr, iters, presimp, mask=mask, ctx=ctx, want_grad=False)
    assert st["integrated_vertices"] == 0
    assert (lab != -1).all()
    mism = float((lab.ravel() != ref["cleaned"]).mean())
    _say(f"  combinatorial-only mismatch vs reference {mism:.5f}")
    assert mism < 0.2


@pytest.mark.parametrize("name", ["labels_48", "smooth_34x33x32", "smooth_presimp_40"])
def test_stage_entry_points_match_oracle(ctx, name):
    f, (err, gthr, iters, presimp), ref = load_ondemand(name)
    # cleanup as a function of its input: the reference's own integrated labels
    cleaned = labels.clean_labels(f, ref["integrated"].reshape(f.shape), ctx=ctx).ravel()
    assert (cleaned == ref["cleaned"]).all(), int((cleaned != ref["cleaned"]).sum())
    r = labels.restriction_labels(ref["cleaned"].reshape(f.shape), ctx=ctx).ravel()
    assert (r == ref["restriction"]).all()
    g = labels.constrained_grad(f, ref["restriction"], ctx=ctx).ravel()
    assert (g == ref["grad"]).all(), int((g != ref["grad"]).sum())


@pytest.mark.parametrize("shape", [(23, 29, 37), (20, 40, 64), (9, 17, 5)])
def test_clean_labels_matches_oracle_on_fragmented_labels(ctx, shape):
    # many small same-label components, most of them crossing the 8x8x4
    # tiles the GPU joins locally first, plus -1 (unlabelled) vertices and
    # tied values (the total order then falls back to vertex ids)
    rng = np.random.default_rng(sum(shape))
    f = np.round(rng.standard_normal(shape) * 4.0).astype(np.float32)
    Z, Y, X = shape
    zz, yy, xx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    lab = ((xx // 5) + 7 * (yy // 3) + 31 * (zz // 2)).astype(np.int32) % 11
    flip = rng.random(shape) < 0.15
    lab[flip] = rng.integers(-1, 11, size=int(flip.sum())).astype(np.int32)
    _say(f"clean fragmented {shape}")
    got = labels.clean_labels(f, lab, ctx=ctx).ravel()
    expect = oracle_lib.clean(f, lab).ravel()
    assert (got == expect).all(), int((got != expect).sum())


@pytest.mark.parametrize("sign", [-1, 1])
@pytest.mark.parametrize("name", ["labels_48", "smooth_34x33x32"])
def test_integrate_labels_matches_oracle(ctx, name, sign):
    # msceer_integrate_labels on the certain regions of f (downhill) or of -f
    # (uphill) against the FP64 oracle integrator; pass 1 depends on the
    # visiting order, so a small mismatch is allowed.
    f, (err, gthr, iters, _), _ = load_ondemand(name)
    src = f if sign < 0 else (-f).astype(np.float32)
    term, seeds, _ = oracle_lib.certain_regions(src, oracle_lib.lstar_pairs(src), 0.0)
    _say(f"integrate_labels {name} sign {sign} {f.shape}")
    lab, st = labels.integrate_labels(f, term.reshape(f.shape), seeds.reshape(f.shape).astype(np.uint8), err, gthr,
                                      iters, sign, ctx=ctx)
    ref, _ = oracle_lib.integrate_signed(f, term, seeds, err, gthr, iters, sign=sign, mode=1)
    mism = float((lab.ravel() != ref).mean())
    _say(f"  mismatch vs FP64 oracle {mism:.5f}, {st}")
    assert mism <= 1e-3
    t = term.ravel()
    assert (lab.ravel()[t != -1] == t[t != -1]).all()  # terminals keep their labels
    assert float((lab == -1).mean()) < 0.01


def test_integrate_labels_uphill_is_downhill_of_the_negated_field(ctx):
    f, (err, gthr, iters, _), _ = load_ondemand("smooth_34x33x32")
    neg = (-f).astype(np.float32)
    term, seeds, _ = oracle_lib.certain_regions(neg, oracle_lib.lstar_pairs(neg), 0.0)
    t, sd = term.reshape(f.shape), seeds.reshape(f.shape).astype(np.uint8)
    up, _ = labels.integrate_labels(f, t, sd, err, gthr, iters, 1, ctx=ctx)
    down, _ = labels.integrate_labels(neg, t, sd, err, gthr, iters, -1, ctx=ctx)
    # every product and sum only changes sign, and the labels do not depend
    # on the order streamlines run in: identical
    assert (up == down).all(), int((up != down).sum())


@pytest.mark.parametrize("name", ["gauss_44x36x40", "smooth_presimp_40"])
def test_labels_are_the_same_from_run_to_run(ctx, name):
    # pass 1's levels are synchronous and pass 3 reads labels fixed before it
    # starts: every stage gives the same result on every call, and the
    # integrator's work counts are the same too
    f, (err, gthr, iters, presimp), _ = load_ondemand(name)
    keys = ("traced_fp32", "reintegrated_pass3", "escalated_fp64", "reintegrated_fp64", "over_budget_fp32",
            "over_budget_fp64")
    for stop in ("integrated", None):
        runs = [labels.ondemand_asc(f, err, gthr, iters, presimp, ctx=ctx, stop=stop) for _ in range(3)]
        _say(f"reproducible {name} stop={stop}: {[{k: r[2][k] for k in keys} for r in runs]}")
        for g, lab, st in runs[1:]:
            assert (lab == runs[0][1]).all(), int((lab != runs[0][1]).sum())
            if stop is None:
                assert (g == runs[0][0]).all()
            assert {k: st[k] for k in keys} == {k: runs[0][2][k] for k in keys}


def test_integrate_labels_rejects_bad_arguments(ctx):
    f = np.zeros((4, 5, 6), np.float32)
    t = np.full(f.shape, -1, np.int32)
    for kw in ({"sign": 0}, {"sign": 2}, {"err_thr": 0.0}, {"max_iters": 0}):
        with pytest.raises(native.MsceerError):
            labels.integrate_labels(f, t, ctx=ctx, **kw)


@pytest.mark.parametrize("name", ["labels_48", "smooth_34x33x32", "gauss_44x36x40"])
def test_max_boundary_bit_exact(ctx, name):
    # onto the reference's MIN restriction: the reference's ascending labels,
    # then descending labels from the oracle integrator (uphill, terminals from -f)
    f, (err, gthr, iters, _), ref = load_ondemand(name)
    neg = (-f).astype(np.float32)
    term, seeds, _ = oracle_lib.certain_regions(neg, oracle_lib.lstar_pairs(neg), 0.0)
    dsc, _ = oracle_lib.integrate_signed(f, term, seeds, err, gthr, iters, sign=1, mode=1)
    for lab in (ref["cleaned"], dsc):
        _say(f"max boundary {name} {f.shape}")
        got = labels.restriction_add_max(f, lab.reshape(f.shape), ref["restriction"], ctx=ctx)
        expect = oracle_lib.max_boundary(f, lab, ref["restriction"])
        assert (got == expect).all(), int((got != expect).sum())
        assert int((expect > 1).sum()) > 0
