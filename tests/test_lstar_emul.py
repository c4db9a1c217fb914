"""The lower-star kernel's per-lane logic, run on the CPU (tests/emul), against
the oracle: the steepest pass on every golden case and seeded fields, and the
constrained pass with random restriction labels."""
import numpy as np
import pytest

import emul_lib
import oracle_lib
from golden import lstar_case_names, load_lstar
from msceer_amd import synth


@pytest.mark.parametrize("name", lstar_case_names())
def test_emulated_kernel_matches_golden_pairs(name):
    f, ref = load_lstar(name)
    expect = emul_lib.expected_prepass(ref & 0x1F, f.shape)
    got = emul_lib.lstar_prepass(f)
    assert (got == expect).all(), int((got != expect).sum())


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_emulated_constrained_pass_matches_oracle(seed):
    s = synth.Stream(seed)
    shape = (9, 14, 23)
    f = s.normal(int(np.prod(shape))).reshape(shape).astype(np.float32)
    if seed == 3:
        f = np.floor(f * 2).astype(np.float32)  # ties
    Z, Y, X = shape
    C = (2 * X - 1) * (2 * Y - 1) * (2 * Z - 1)
    lab = np.floor(s.uniform(C) * 4).astype(np.int8)  # labels 0..3 (key collisions included)
    ref = oracle_lib.flags(oracle_lib.lstar_pairs(f, lab), shape)
    got = emul_lib.lstar_prepass(f, lab)
    expect = emul_lib.expected_prepass(ref, shape)
    assert (got == expect).all(), int((got != expect).sum())


@pytest.mark.parametrize("kind", ["smooth", "noise", "ties", "plateau"])
def test_closed_form_product_stars_equal_general_expansion(kind):
    # octants, quadrants, single edges and minima: the closed form must write
    # exactly the bytes the general loops give, on exactly the lower star,
    # with tied values included (plateaus and rounded noise make many)
    shape = (40, 44, 48)
    s = synth.Stream(77)
    if kind == "smooth":
        f = synth.smoothed_noise(shape, 77, 2.0)
    elif kind == "noise":
        f = s.normal(int(np.prod(shape))).reshape(shape).astype(np.float32)
    elif kind == "ties":
        f = np.floor(s.normal(int(np.prod(shape))).reshape(shape) * 2).astype(np.float32)
    else:
        f = np.floor(synth.smoothed_noise(shape, 78, 3.0) * 20).astype(np.float32)
    taken, differ = emul_lib.check_product(f)
    assert taken > 0.2 * (shape[0] - 2) * (shape[1] - 2) * (shape[2] - 2)
    assert differ == 0, (taken, differ)
