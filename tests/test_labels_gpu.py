"""GPU accurate labelling (msceer_ondemand_asc) against the reference goldens
and the CPU oracle.

Deterministic stages are bit-exact: certain regions, and the conforming .grad
given the labels.  The integrated labels are compared as a mismatch rate with
the FP64 oracle (pass 1 of the reference depends on its visiting order, and
the GPU traces in FP32 first): tolerance 1% of vertices on these small
volumes (the larger-volume target is reported by bench.py).
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
    assert mism < 0.01
    # the .grad is a pure function of the labels: bit-exact
    r = oracle_lib.min_boundary(lab, f.shape)
    expect = oracle_lib.flags(oracle_lib.ldir(oracle_lib.lstar_pairs(f, r), f.shape), f.shape)
    assert (grad == expect).all(), int((grad != expect).sum())
    gradcheck.check_consistency(grad, X, Y, Z)


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
