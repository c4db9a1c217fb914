"""The oracle's descending 3-manifold branch (needDSC2) against the
reference's own run (tests/golden/dsc2, written by make_dsc2_golden.py):
certain regions, integrated and cleaned labels as partitions (the region
ids are the reference's own numbering), restriction labels and the
conforming .grad bit-exact, alone and after needASC2 (which shares the
gradient field, zeroed at the ascending seeds too)."""
import hashlib

import numpy as np
import pytest

import oracle_lib
from golden import dsc2_case_names, load_dsc2, same_partition


@pytest.mark.parametrize("tag", dsc2_case_names())
def test_reference_order_is_bit_exact(tag):
    f, (err, gthr, iters, presimp), flags, ref, grad_sha = load_dsc2(tag)
    g, la, ld, r, st = oracle_lib.ondemand2(f, err, gthr, iters, presimp, *flags, mode=0)
    assert same_partition(ld, ref["cleaned"])
    assert (r == ref["restriction"]).all(), int((r != ref["restriction"]).sum())
    assert hashlib.sha256(g.tobytes()).hexdigest() == grad_sha


@pytest.mark.parametrize("tag", ["smooth_34x33x32_0001", "labels_48_0011"])
def test_fp64_mode_stays_close_to_the_reference(tag):
    """Mode 1 (every quantity in FP64, the GPU's arithmetic for the
    re-integrations) moves few labels from the reference's mixed arithmetic."""
    f, (err, gthr, iters, presimp), flags, ref, _ = load_dsc2(tag)
    _, _, l0, _, _ = oracle_lib.ondemand2(f, err, gthr, iters, presimp, *flags, mode=0)
    _, _, l1, _, _ = oracle_lib.ondemand2(f, err, gthr, iters, presimp, *flags, mode=1)
    assert (l0 != l1).mean() <= 5e-3
