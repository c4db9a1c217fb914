"""The GPU's 1-manifold algorithm (speculation over windows, union-find over
footprints, in-order replay of components, commit before a cut), built for
the host from the kernels' own header and checked against the oracle's
one-thread saddle loops in FP64: equal restriction labels for every lane
order of a launch and for windows small enough to cut often."""
import numpy as np
import pytest

import emul_lib
import oracle_lib
from golden import load_lines
from msceer_amd import synth


def _check(f, windows=(1 << 20, 61), orders=(0, 1, 2), err=1e-3, gthr=1e-6, iters=4096):
    base = oracle_lib.lstar_pairs(f)
    field = oracle_lib.gradient_field(f)
    for asc in (1, 0):
        ref, ost = oracle_lib.saddle_lines(f, base, asc, err=err, gthr=gthr, iters=iters, mode=1)
        for window in windows:
            for order in orders:
                r, st = emul_lib.lines(f, base, field, asc, window=window, order=order, err=err, gthr=gthr,
                                       iters=iters)
                assert (r == ref).all(), (asc, window, order, int((r != ref).sum()), st.tolist())
                assert st[0] == ost[0]


@pytest.mark.parametrize("name", ["labels_48", "gauss_44x36x40", "smooth_34x33x32", "budget_24x30x20"])
def test_emulated_lines_equal_the_oracle_on_the_golden_volumes(name):
    f, (err, gthr, iters, presimp), _, _ = load_lines(name, (1, 1, 0))
    _check(f, err=err, gthr=gthr, iters=iters)


def test_emulated_lines_equal_the_oracle_on_a_larger_volume():
    _check(synth.smoothed_noise((72, 80, 96), 303, 2.0), windows=(1 << 20, 97), orders=(0, 2))
