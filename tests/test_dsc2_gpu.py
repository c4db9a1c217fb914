"""GPU descending 3-manifolds (needDSC2 through msceer_ondemand2) against the
reference's own run (tests/golden/dsc2) and the CPU oracle.

As for the ascending labels: the certain regions follow the reference
operation for operation and match its regions exactly (as a partition: the
reference numbers regions, the GPU names them by a vertex); the integrated
and cleaned labels are compared with the FP64 oracle as mismatch rates (from
the reference's visiting order or from every vertex traced alone, at most
0.1 % of the vertices or as far as those two are apart, and at most 0.25 %
from the former); the restriction labels and the conforming .grad
are a pure function of the labels and are bit-exact given them.  Each case
runs alone (needDSC2) and after needASC2, which shares the gradient field.
"""
import ctypes
import sys

import numpy as np
import pytest

import gradcheck
import oracle_lib
from golden import dsc2_case_names, load_dsc2, same_partition
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


def _run(ctx, f, params, flags, stop=None):
    err, gthr, iters, presimp = params
    asc1, dsc1, asc2, dsc2 = flags
    Z, Y, X = f.shape
    need = ((native.NEED_ASC1 if asc1 else 0) | (native.NEED_DSC1 if dsc1 else 0) |
            (native.NEED_ASC2 if asc2 else 0) | (native.NEED_DSC2 if dsc2 else 0))
    la = np.full(f.size, -7, np.int32)
    ld = np.full(f.size, -7, np.int32)
    st = native.LabelStats()
    fl = {None: 0, "regions": native.STOP_REGIONS, "integrated": native.STOP_INTEGRATED}[stop]
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    f = np.ascontiguousarray(f, np.float32)
    grad = None if stop else np.empty(oracle_lib.ncells(X, Y, Z), np.uint8)
    restr = None if stop else np.empty(oracle_lib.ncells(X, Y, Z), np.int8)
    rc = native.lib().msceer_ondemand2(ctx.handle, ptr(f), X, Y, Z, err, gthr, iters, presimp, need, None,
                                       None if grad is None else ptr(grad), ptr(la) if asc2 else None,
                                       ptr(ld) if dsc2 else None,
                                       None if restr is None else ptr(restr), ctypes.byref(st), fl)
    native.check(rc, "msceer_ondemand2")
    return grad, la, ld, restr, st.as_dict()


def _check_bars(mism, mexact, spread):
    """mism: GPU vs the reference's visiting order (FP64); mexact: GPU vs
    every vertex traced alone in FP64; spread: those two integrations apart."""
    _say(f"  reference order vs traced alone {spread:.5f}")
    assert min(mism, mexact) <= max(1e-3, spread)
    assert mism <= 2.5e-3


@pytest.mark.parametrize("tag", dsc2_case_names())
def test_certain_regions_match_the_reference(ctx, tag):
    f, params, flags, ref, _ = load_dsc2(tag)
    _say(f"dsc2 regions {tag} {f.shape}")
    _, _, ld, _, st = _run(ctx, f, params, flags, stop="regions")
    c = ref["certain"].ravel()
    assert ((ld == -1) == (c == -1)).all(), int(((ld == -1) != (c == -1)).sum())
    assert same_partition(ld, c)
    # regions are named by a maximum's highest vertex, labelled by itself
    lab = np.unique(ld[ld >= 0])
    assert (ld[lab] == lab).all()
    assert st["regions_dsc"] == len(lab), (st["regions_dsc"], len(lab))


@pytest.mark.parametrize("tag", dsc2_case_names())
def test_labels_restriction_and_grad(ctx, tag):
    f, params, flags, ref, grad_sha = load_dsc2(tag)
    Z, Y, X = f.shape
    _say(f"dsc2 full {tag} {f.shape}")
    grad, la, ld, restr, st = _run(ctx, f, params, flags)
    _, la64, ld64, _, _ = oracle_lib.ondemand2(f, *params, *flags, mode=1)
    ldx = oracle_lib.ondemand2(f, *params, *flags, mode=2)[2].ravel()
    mism = float((ld != ld64.ravel()).mean())
    mexact = float((ld != ldx).mean())
    _say(f"  dsc mismatch vs FP64 oracle {mism:.5f}, vs every vertex traced alone {mexact:.5f}; "
         f"maxima {st['maxima']} saddles2 {st['saddles2']} regions {st['regions_dsc']} "
         f"sweeps {st['certain_sweeps_dsc']} ms_dsc {st['ms_dsc']:.3f} ms_total {st['ms_total']:.3f}")
    # Vertices left unlabelled (-1) are part of the comparison: uphill
    # streamlines that end at a maximum on the domain boundary (no critical
    # hex there), or after needASC2 at a zeroed ascending seed, stay -1 in the
    # reference too.  Along the domain's edges the field is one-dimensional
    # and nearly flat, and which of those vertices the reference's pass 1
    # labels by path compression depends on its visiting order: on the
    # Gaussian-mixture case its own order differs from every vertex traced
    # alone in FP64 on 0.15 % of the vertices, the GPU on 0.03 %.  (With a
    # small step budget, as in the budget cases, tracing alone differs in
    # turn: a path-compressed streamline ends early on a labelled vertex.)
    # Hence: no farther from the reference order or from the order-free
    # integration than 0.1 % or than those two are from each other, and at
    # most 0.25 % from the reference order.
    _check_bars(mism, mexact, float((ld64.ravel() != ldx).mean()))
    if flags[2]:
        assert float((la != la64.ravel()).mean()) <= 1e-3
    # restriction labels and .grad: a pure function of the labels
    r = oracle_lib.min_boundary(la, f.shape) if flags[2] else np.zeros(oracle_lib.ncells(X, Y, Z), np.int8)
    r = oracle_lib.max_boundary(f, ld, r)
    assert (restr == r).all(), int((restr != r).sum())
    expect = oracle_lib.flags(oracle_lib.ldir(oracle_lib.lstar_pairs(f, r), f.shape), f.shape)
    assert (grad == expect).all(), int((grad != expect).sum())
    gradcheck.check_consistency(grad, X, Y, Z)


def test_stopped_stages_are_monotone(ctx):
    """Certain vertices keep their region through integration; the cleanup
    makes no new region."""
    f, params, flags, _, _ = load_dsc2("smooth_34x33x32_0011")
    _, la0, l0, _, _ = _run(ctx, f, params, flags, stop="regions")
    _, la1, l1, _, _ = _run(ctx, f, params, flags, stop="integrated")
    _, _, l2, _, _ = _run(ctx, f, params, flags)
    c = l0 >= 0
    assert (l1[c] == l0[c]).all()
    assert (l1 >= 0).sum() > c.sum()
    assert set(np.unique(l2)) <= set(np.unique(l1)) | {-1}


def test_the_ascending_labels_do_not_change(ctx):
    """needDSC2 leaves the ascending labels and their stats as needASC2 alone
    gives them."""
    f, params, _, _, _ = load_dsc2("labels_48_0011")
    _, la_alone, _, _, _ = _run(ctx, f, params, (0, 0, 1, 0))
    _, la_both, _, _, _ = _run(ctx, f, params, (0, 0, 1, 1))
    assert (la_alone == la_both).all()


def test_labels_are_the_same_from_run_to_run(ctx):
    f, params, flags, _, _ = load_dsc2("gauss_44x36x40_0001")
    _, _, a, _, _ = _run(ctx, f, params, flags)
    _, _, b, _, _ = _run(ctx, f, params, flags)
    assert (a == b).all()


def test_wrapper_and_argument_checks(ctx):
    f, (err, gthr, iters, presimp), _, _, _ = load_dsc2("budget_24x30x20_0001")
    g, la, r, st, ld = labels.ondemand(f, need_asc2=False, need_dsc2=True, err_thr=err, grad_thr=gthr,
                                       max_iters=iters, presimp=presimp, ctx=ctx, with_dsc=True)
    assert la is None and ld.shape == f.shape and (ld >= 0).all()
    Z, Y, X = f.shape
    out = np.empty(f.size, np.int32)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    lib = native.lib()
    # labels_dsc_out without needDSC2
    assert lib.msceer_ondemand2(ctx.handle, ptr(f), X, Y, Z, err, gthr, iters, presimp, native.NEED_ASC2, None,
                                None, None, ptr(out), None, None, 0) == native.EINVAL
    # an unknown need bit
    assert lib.msceer_ondemand2(ctx.handle, ptr(f), X, Y, Z, err, gthr, iters, presimp, 0x10, None,
                                None, None, None, None, None, 0) == native.EINVAL
    # ascending labels without needASC2
    assert lib.msceer_ondemand2(ctx.handle, ptr(f), X, Y, Z, err, gthr, iters, presimp, native.NEED_DSC2, None,
                                None, ptr(out), None, None, None, 0) == native.EINVAL


def test_larger_volume_against_the_oracle(ctx):
    """96^3 mixture of 40 Gaussians of either sign, needASC2 + needDSC2: the
    descending labels within the bars above of the FP64 oracle, restriction
    and .grad bit-exact given the labels."""
    from msceer_amd import synth
    f = (synth.gaussian_mixture((96, 96, 96), 77, 40, 3.0, 9.0, -1.0, 1.0)).astype(np.float32)
    params = (1e-3, 1e-6, 4096, 0.0)
    flags = (0, 0, 1, 1)
    Z, Y, X = f.shape
    _say("dsc2 96^3 gaussian mixture")
    grad, la, ld, restr, st = _run(ctx, f, params, flags)
    ld64 = oracle_lib.ondemand2(f, *params, *flags, mode=1)[2].ravel()
    ldx = oracle_lib.ondemand2(f, *params, *flags, mode=2)[2].ravel()
    mism, mexact = float((ld != ld64).mean()), float((ld != ldx).mean())
    _say(f"  dsc mismatch vs FP64 oracle {mism:.5f}, vs every vertex traced alone {mexact:.5f}; "
         f"maxima {st['maxima']} regions {st['regions_dsc']} ms_dsc {st['ms_dsc']:.3f}")
    _check_bars(mism, mexact, float((ld64 != ldx).mean()))
    r = oracle_lib.max_boundary(f, ld, oracle_lib.min_boundary(la, f.shape))
    assert (restr == r).all(), int((restr != r).sum())
    expect = oracle_lib.flags(oracle_lib.ldir(oracle_lib.lstar_pairs(f, r), f.shape), f.shape)
    assert (grad == expect).all(), int((grad != expect).sum())


@pytest.mark.parametrize("flags", [(1, 1, 0, 1), (1, 1, 1, 1)])
def test_one_manifolds_after_the_descending_labels(ctx, flags):
    """needASC1 / needDSC1 with needDSC2: the digitized 1-manifolds follow
    the FP64 field zeroed at the seeds of every labelling that ran (the
    descending branch's field): cells 4 and 5 equal the FP64 oracle's."""
    f, params, _, _, _ = load_dsc2("smooth_34x33x32_0001")
    Z, Y, X = f.shape
    _say(f"dsc2 with 1-manifolds {flags}")
    grad, la, ld, restr, st = _run(ctx, f, params, flags)
    _, ola, old, o_r, ost = oracle_lib.ondemand2(f, *params, *flags, mode=1)
    for t in (4, 5):
        assert ((restr == t) == (o_r == t)).all(), (t, int(((restr == t) != (o_r == t)).sum()))
    assert float((ld != old.ravel()).mean()) <= 1e-3
    if flags[2]:
        assert float((la != ola.ravel()).mean()) <= 1e-3
    assert st["lines_asc"] >= ost[5] and st["lines_dsc"] >= ost[13]
    expect = oracle_lib.flags(oracle_lib.ldir(oracle_lib.lstar_pairs(f, restr), f.shape), f.shape)
    assert (grad == expect).all(), int((grad != expect).sum())
    gradcheck.check_consistency(grad, X, Y, Z)


def test_labelling_volume_256(ctx):
    """The 256^3 labelling volume (200 Gaussians + sine, seed 303), needASC2 +
    needDSC2 beside each other: the descending certain regions bit-exact
    against the oracle's at full size, kept through the integration; the
    restriction labels and the .grad a pure function of the labels."""
    import ctypes as C
    from golden import C3_PARAMS
    from msceer_amd import synth
    f = synth.labels_field(303, 256).astype(np.float32)
    Z, Y, X = f.shape
    flags = (0, 0, 1, 1)
    _say("dsc2 256^3: certain regions")
    _, _, l0, _, _ = _run(ctx, f, C3_PARAMS, flags, stop="regions")
    base = oracle_lib.lstar_pairs(f)
    lab = np.full(f.size, -1, np.int32)
    seeds = np.zeros(f.size, np.int8)
    h = oracle_lib.lib()
    h.oracle_certain_regions_dsc.argtypes = [C.c_void_p, C.c_longlong, C.c_longlong, C.c_longlong, C.c_void_p,
                                             C.c_float, C.c_void_p, C.c_void_p]
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    h.oracle_certain_regions_dsc(p(f), X, Y, Z, p(base), float(C3_PARAMS[3]), p(lab), p(seeds))
    assert (l0 == lab).all(), int((l0 != lab).sum())
    _say("dsc2 256^3: integrated")
    _, _, l1, _, _ = _run(ctx, f, C3_PARAMS, flags, stop="integrated")
    c = lab >= 0
    assert (l1[c] == lab[c]).all()
    _say("dsc2 256^3: full")
    grad, la, ld, restr, st = _run(ctx, f, C3_PARAMS, flags)
    _say(f"  maxima {st['maxima']} regions {st['regions_dsc']} ms_dsc {st['ms_dsc']:.3f} "
         f"ms_total {st['ms_total']:.3f} unlabelled {float((ld < 0).mean()):.4f}")
    r = oracle_lib.max_boundary(f, ld, oracle_lib.min_boundary(la, f.shape))
    assert (restr == r).all(), int((restr != r).sum())
    expect = oracle_lib.flags(oracle_lib.ldir(oracle_lib.lstar_pairs(f, r), f.shape), f.shape)
    assert (grad == expect).all(), int((grad != expect).sum())
