"""NumericIntegratorExpandingRegionStop on the GPU
(msceer_expanding_region_stop) against the oracle's order-free FP64 form
(the same computation: every vertex traced alone) and against the
reference's own class run with one thread, which the oracle reproduces bit
for bit in its reference order."""
import sys

import numpy as np
import pytest

import oracle_lib
from golden import load_ers, ondemand_case_names
from msceer_amd import labels, native, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name", ondemand_case_names())
def test_against_oracle_and_reference(ctx, name):
    f, (err, gthr, iters), ref, extrema, spread = load_ers(name)
    lab, st = labels.expanding_region_stop(f, err, gthr, iters, ctx=ctx)
    olab, ost = oracle_lib.expanding_region_stop(f, err, gthr, iters, 2)
    d_oracle = float((lab != olab).mean())
    d_ref = float((lab != ref).mean())
    print(f"{name}: stats {st.tolist()} oracle {ost.tolist()}; vs oracle {100 * d_oracle:.4f} %, vs one-thread "
          f"reference {100 * d_ref:.3f} % (its 8-thread runs {[round(100 * x, 3) for x in spread]} %)", flush=True)
    sys.stdout.flush()
    assert st[0] == extrema and st[0] == ost[0] and st[1] == ost[1]
    assert d_oracle <= 1e-4
    assert d_ref <= 0.025


@pytest.mark.parametrize("shape,seed", [((64, 72, 80), 61), ((40, 100, 30), 62)])
def test_larger_volumes_against_oracle(ctx, shape, seed):
    f = synth.smoothed_noise(shape, seed, 2.5).astype(np.float32)
    lab, st = labels.expanding_region_stop(f, 1e-3, 1e-6, 4096, ctx=ctx)
    olab, ost = oracle_lib.expanding_region_stop(f, 1e-3, 1e-6, 4096, 2)
    d = float((lab != olab).mean())
    print(f"{shape}: stats {st.tolist()} vs oracle {100 * d:.4f} %", flush=True)
    assert st[0] == ost[0] and st[1] == ost[1]
    assert d <= 1e-4
