"""The CPU oracle of the lower-star gradient, pinned before it is trusted.

* golden vectors: .grad bytes written by the reference's own classes
  (tests/golden/make_lstar_golden.py) -- bit-exact, every case;
* max/min vertex labels against a brute-force argmax/argmin in numpy;
* gradient invariants (consistency, Euler characteristic, ldir closed form).
"""
import numpy as np
import pytest

import gradcheck
import oracle_lib
from golden import lstar_case_names, load_lstar


@pytest.mark.parametrize("name", lstar_case_names())
def test_oracle_matches_reference_golden(name):
    f, ref = load_lstar(name)
    got = oracle_lib.steepest_grad(f)
    bad = np.nonzero(got != ref)[0]
    assert bad.size == 0, f"{name}: {bad.size} bytes differ, first at cell {bad[:5]}"


def _brute_maxmin(f):
    Z, Y, X = f.shape
    Mx, My, Mz = 2 * X - 1, 2 * Y - 1, 2 * Z - 1
    flat = f.ravel()
    mx = np.empty(Mx * My * Mz, np.uint8)
    mn = np.empty_like(mx)
    c = 0
    for cz in range(Mz):
        for cy in range(My):
            for cx in range(Mx):
                best_hi = best_lo = None
                for dz in ((-1, 1) if cz & 1 else (0,)):
                    for dy in ((-1, 1) if cy & 1 else (0,)):
                        for dx in ((-1, 1) if cx & 1 else (0,)):
                            v = (cx + dx) // 2 + X * ((cy + dy) // 2 + Y * ((cz + dz) // 2))
                            key = (float(flat[v]), v)
                            tid = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)
                            if best_hi is None or key > best_hi[0]:
                                best_hi = (key, tid)
                            if best_lo is None or key < best_lo[0]:
                                best_lo = (key, tid)
                mx[c], mn[c] = best_hi[1], best_lo[1]
                c += 1
    return mx, mn


@pytest.mark.parametrize("name", ["noise_17x9x5", "ties_10x11x12", "signed_zeros_9x8x7", "tiny_5x2x3"])
def test_maxmin_brute_force(name):
    f, _ = load_lstar(name)
    mx, mn = oracle_lib.maxmin(f)
    bmx, bmn = _brute_maxmin(f)
    assert (mx == bmx).all() and (mn == bmn).all()


@pytest.mark.parametrize("name", lstar_case_names())
def test_golden_invariants(name):
    f, ref = load_lstar(name)
    Z, Y, X = f.shape
    gradcheck.check_consistency(ref, X, Y, Z)
    # flag bit is boundaryValue != 0
    _, (cx, cy, cz) = gradcheck.cell_dims(X, Y, Z)
    bv = ((cx == 0) | (cx == 2 * X - 2)) | ((cy == 0) | (cy == 2 * Y - 2)) | ((cz == 0) | (cz == 2 * Z - 2))
    assert (((ref >> 1) & 1) == bv.ravel()).all()


@pytest.mark.parametrize("name", ["c1_gauss64", "noise_48x97x31", "smooth_60x50x40", "plateau_33x29x21"])
def test_ldir_closed_form(name):
    f, ref = load_lstar(name)
    Z, Y, X = f.shape
    assert ((ref >> 5) == gradcheck.ldir_reachability(ref, X, Y, Z)).all()


def test_oracle_ldir_recomputes_from_pairs():
    f, ref = load_lstar("smooth_60x50x40")
    pairs = ref & 0x1D
    g = oracle_lib.flags(oracle_lib.ldir(pairs, f.shape), f.shape)
    assert (g == ref).all()


def test_constant_field_criticals_only_on_boundary_strata():
    # ties reduce the order to the vertex id: one critical cell per boundary
    # stratum subset (27 for any box of at least 2x2x2)
    f, ref = load_lstar("const_8x7x6")
    crit = ((ref >> 2) & 7) == 7
    assert crit.sum() == 27
    assert ((ref[0] >> 2) & 7) == 7  # vertex 0 is the global minimum


def test_planted_tie_break_mistake_is_caught():
    # flipping the tie-break of equal values (a plausible porting mistake)
    # must change the .grad of a field with plateaus
    f, ref = load_lstar("ties_10x11x12")
    flipped = f[::-1, ::-1, ::-1].copy()
    g = oracle_lib.steepest_grad(flipped)
    assert (g[::-1] != ref).any()
