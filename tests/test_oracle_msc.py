"""The oracle's reading of a .grad as extractmsc reads it (nodes, base basins,
base mountains) against the reference's own Morse-Smale complex class run on
the reference .grad of the lower-star goldens (tests/golden/msc, written by
make_msc_golden.py from MorseSmaleComplexBasic::ComputeFromGrad and
fillGeometry), and against what the topology fixes: the critical cells'
alternating sum is the Euler characteristic of a box (1), every vertex lies
in some basin, and the mountain labels cover exactly the vertices that own a
hex, less those on chains that end at a boundary quad."""
import numpy as np
import pytest

import oracle_lib
from golden import load_msc, msc_case_names


@pytest.mark.parametrize("name", msc_case_names())
def test_oracle_matches_reference_msc(name):
    f, grad, basins, mounts, counts = load_msc(name)
    b, m, c = oracle_lib.msc_base(grad, f.shape)
    assert (c == counts).all(), (c, counts)
    assert (b == basins).all(), int((b != basins).sum())
    assert (m == mounts).all(), int((m != mounts).sum())


@pytest.mark.parametrize("name", msc_case_names())
def test_euler_characteristic_and_coverage(name):
    f, grad, basins, mounts, counts = load_msc(name)
    b, m, c = oracle_lib.msc_base(grad, f.shape)
    assert c[0] - c[1] + c[2] - c[3] == 1
    assert (b >= 0).all() and b.max() < c.sum()
    Z, Y, X = f.shape
    assert (m[Z - 1:, :, :] == -1).all() and (m[:, Y - 1:, :] == -1).all() and (m[:, :, X - 1:] == -1).all()


def test_a_planted_mistake_is_caught():
    """Swapping a vertex's pair direction breaks the match with the reference."""
    f, grad, basins, mounts, counts = load_msc("smooth_60x50x40")
    g = grad.copy().reshape(2 * f.shape[0] - 1, 2 * f.shape[1] - 1, 2 * f.shape[2] - 1)
    # a non-critical interior vertex paired along +x: point it along -x instead
    zz, yy, xx = np.nonzero(((g[::2, ::2, ::2] >> 2) & 7) == 0)
    k = [i for i in range(len(xx)) if 2 < xx[i] < f.shape[2] - 3][0]
    cz, cy, cx = 2 * zz[k], 2 * yy[k], 2 * xx[k]
    g[cz, cy, cx] = (g[cz, cy, cx] & ~np.uint8(0x1C)) | np.uint8(1 << 2)
    b, _, _ = oracle_lib.msc_base(g.ravel(), f.shape)
    assert (b != basins).any()
