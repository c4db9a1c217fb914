"""The oracle's NumericIntegratorExpandingRegionStop against the reference's
own class run with one thread (tests/golden/ers, written by
make_ers_golden.py through oracle/refdrv/ref_ers.cpp): the reference
arithmetic (mode 0) and the same order in FP64 (mode 1) are bit-exact on
every case, including a step budget of 60; the order-free FP64 form (mode 2,
what the GPU computes) stays within a few times the reference's own
thread-schedule spread."""
import numpy as np
import pytest

import oracle_lib
from golden import load_ers, ondemand_case_names


@pytest.mark.parametrize("name", ondemand_case_names())
@pytest.mark.parametrize("mode", [0, 1])
def test_reference_order_is_bit_exact(name, mode):
    f, (err, gthr, iters), ref, extrema, _ = load_ers(name)
    lab, st = oracle_lib.expanding_region_stop(f, err, gthr, iters, mode)
    assert st[0] == extrema
    assert (lab == ref).all(), int((lab != ref).sum())


@pytest.mark.parametrize("name", ondemand_case_names())
def test_order_free_form_near_the_reference(name):
    f, (err, gthr, iters), ref, extrema, spread = load_ers(name)
    lab, st = oracle_lib.expanding_region_stop(f, err, gthr, iters, 2)
    mism = float((lab != ref).mean())
    print(f"{name}: order-free vs one-thread reference {100 * mism:.3f} %, 8-thread reference "
          f"{[round(100 * x, 3) for x in spread]} %")
    assert st[0] == extrema
    assert mism <= 0.025


def test_minima_are_their_own_labels_and_corners_are_minima():
    f, (err, gthr, iters), ref, extrema, _ = load_ers("labels_48")
    lab, _ = oracle_lib.expanding_region_stop(f, err, gthr, iters, 2)
    flat = lab.ravel()
    mins = np.unique(flat[flat >= 0])
    assert (flat[mins] == mins).all()
    Z, Y, X = f.shape
    for (z, y, x) in [(0, 0, 0), (Z - 1, 0, 0), (0, Y - 1, X - 1), (Z - 1, Y - 1, X - 1)]:
        v = x + X * (y + Y * z)
        assert flat[v] == v  # a corner has no SameBdry6 neighbour
