"""The CPU oracle of accurate ascending labelling, pinned before it is trusted:
every stage against the reference's own outputs (golden vectors written by
oracle/refdrv/ref_ondemand.cpp with one thread), plus closed forms."""
import numpy as np
import pytest

import gradcheck
import oracle_lib
from golden import canonical, load_ondemand, ondemand_case_names


@pytest.mark.parametrize("name", ondemand_case_names())
def test_oracle_stages_match_reference(name):
    f, (err, gthr, iters, presimp), ref = load_ondemand(name)
    Z, Y, X = f.shape
    base = oracle_lib.lstar_pairs(f)
    cert, seeds, _ = oracle_lib.certain_regions(f, base, presimp)
    assert (cert == ref["certain"]).all(), int((cert != ref["certain"]).sum())
    integ, _ = oracle_lib.integrate(f, cert, seeds, err, gthr, iters, mode=0)
    assert (canonical(integ, f) == ref["integrated"]).all(), int((canonical(integ, f) != ref["integrated"]).sum())
    cleaned = oracle_lib.clean(f, integ)
    assert (cleaned == ref["cleaned"]).all()
    r = oracle_lib.min_boundary(cleaned, f.shape)
    assert (r == ref["restriction"]).all()
    g, lab, r2, st = oracle_lib.ondemand_asc(f, err, gthr, iters, presimp, mode=0)
    assert (g == ref["grad"]).all(), int((g != ref["grad"]).sum())
    assert (lab == ref["cleaned"]).all() and (r2 == ref["restriction"]).all()
    gradcheck.check_consistency(g, X, Y, Z)


def test_gradient_field_closed_form():
    f, _, _ = load_ondemand("smooth_34x33x32")
    g = oracle_lib.gradient_field(f)
    f64 = f.astype(np.float64)
    expect = np.zeros(f.shape + (3,))
    expect[:, :, 1:-1, 0] = 0.5 * (f64[:, :, 2:] - f64[:, :, :-2])
    expect[:, 1:-1, :, 1] = 0.5 * (f64[:, 2:, :] - f64[:, :-2, :])
    expect[1:-1, :, :, 2] = 0.5 * (f64[2:] - f64[:-2])
    assert np.array_equal(g, expect)


def test_fp64_mode_close_to_reference_arithmetic():
    f, (err, gthr, iters, presimp), ref = load_ondemand("labels_48")
    _, lab64, _, _ = oracle_lib.ondemand_asc(f, err, gthr, iters, presimp, mode=1)
    mism = float((lab64 != ref["cleaned"]).mean())
    assert mism < 0.01, mism


def test_budget_exhaustion_is_forced():
    f, (err, gthr, iters, presimp), ref = load_ondemand("budget_24x30x20")
    base = oracle_lib.lstar_pairs(f)
    cert, seeds, _ = oracle_lib.certain_regions(f, base, presimp)
    _, st = oracle_lib.integrate(f, cert, seeds, err, gthr, iters, mode=0)
    assert st[1] > 0  # some streamlines stopped on OVER_MAX_ITERATIONS


def test_conforming_vertices_cover_restricted_lower_stars():
    f, _, ref = load_ondemand("gauss_44x36x40")
    vs = oracle_lib.conforming_vertices(f, ref["restriction"])
    assert vs.size > 0 and (vs % 2 == 0).all()


@pytest.mark.parametrize("name", ["labels_48", "smooth_34x33x32"])
def test_constrained_pass_over_all_vertices_equals_conforming_recompute(name):
    # the GPU recomputes every lower star under the restriction instead of
    # only the affected ones; unaffected stars split exactly as in the
    # steepest pass, so both give the reference .grad
    f, _, ref = load_ondemand(name)
    g = oracle_lib.lstar_pairs(f, ref["restriction"])
    g = oracle_lib.flags(oracle_lib.ldir(g, f.shape), f.shape)
    assert (g == ref["grad"]).all()


@pytest.mark.parametrize("mode", [0, 1])
def test_oracle_uphill_is_downhill_of_the_negated_field(mode):
    # AdaptiveEulerAdvector3D<F,+1> flows up the gradient of f, which is down
    # the gradient of -f; every product and sum only changes sign, so the
    # labels agree exactly (sign -1 is pinned to the reference above).
    f, (err, gthr, iters, _), _ = load_ondemand("smooth_34x33x32")
    neg = (-f).astype(np.float32)
    term, seeds, _ = oracle_lib.certain_regions(neg, oracle_lib.lstar_pairs(neg), 0.0)
    up, st_up = oracle_lib.integrate_signed(f, term, seeds, err, gthr, iters, sign=1, mode=mode)
    down, st_down = oracle_lib.integrate_signed(neg, term, seeds, err, gthr, iters, sign=-1, mode=mode)
    assert (up == down).all(), int((up != down).sum())
    assert (st_up == st_down).all()
    assert float((up == -1).mean()) < 0.01  # streamlines reach the maxima's regions
    # flowing the other way ends elsewhere
    other, _ = oracle_lib.integrate_signed(f, term, seeds, err, gthr, iters, sign=-1, mode=mode)
    assert float((other != up).mean()) > 0.1


@pytest.mark.parametrize("field", ["ramp", "constant"])
def test_oracle_max_boundary_closed_form(field):
    # f increasing with the vertex id (or constant: ties fall back to id order),
    # so every hex's highest vertex is its (+x,+y,+z) corner; labels split at
    # x = k.  Only quads normal to x in the plane x = k - 1 separate hexes
    # whose highest vertices differ in label, so exactly the interior cells
    # with cell x-coordinate 2(k - 1) get 2 (quads, then the edges and vertices
    # under them), and nothing else changes.
    X, Y, Z, k = 9, 7, 6, 4
    f = (np.arange(X * Y * Z, dtype=np.float32) if field == "ramp" else np.zeros(X * Y * Z, np.float32))
    f = f.reshape(Z, Y, X)
    xs = np.arange(X)[None, None, :].repeat(Y, 1).repeat(Z, 0)
    lab = (xs >= k).astype(np.int32)
    r = oracle_lib.max_boundary(f, lab, np.zeros((2 * Z - 1) * (2 * Y - 1) * (2 * X - 1), np.int8))
    cz, cy, cx = np.meshgrid(np.arange(2 * Z - 1), np.arange(2 * Y - 1), np.arange(2 * X - 1), indexing="ij")
    interior = (cx > 0) & (cx < 2 * X - 2) & (cy > 0) & (cy < 2 * Y - 2) & (cz > 0) & (cz < 2 * Z - 2)
    expect = np.where(interior & (cx == 2 * (k - 1)), 2, 0).astype(np.int8).ravel()
    assert (r == expect).all(), int((r != expect).sum())
    # accumulates onto an existing labelling; constant or unassigned labels add nothing
    ones = np.ones_like(expect)
    assert (oracle_lib.max_boundary(f, lab, ones) == ones + expect).all()
    assert (oracle_lib.max_boundary(f, np.zeros_like(lab), ones) == ones).all()
    assert (oracle_lib.max_boundary(f, np.full_like(lab, -1), ones) == ones).all()


def test_fp64_mode_equals_reference_arithmetic_where_every_quantity_is_exact():
    # f = x: the central-difference gradient is (1, 0, 0) inside, the step is
    # capped at 0.5 / |g| = 0.5 voxel from the start, positions stay on
    # multiples of 0.25 and every interpolation weight is in {0, 1/4, 1/2, 3/4}:
    # no quantity rounds in FP32 or FP64, so both arithmetic modes and the
    # exact-destination mode give the same labels bit for bit
    Z, Y, X = 7, 9, 23
    f = np.broadcast_to(np.arange(X, dtype=np.float32)[None, None, :], (Z, Y, X)).copy()
    labs = [oracle_lib.ondemand_asc(f, 1e-3, 1e-6, 4096, 0.0, mode=m)[1] for m in (0, 1, 2)]
    assert (labs[0] == labs[1]).all() and (labs[1] == labs[2]).all()
    assert (labs[0] >= 0).all()


@pytest.mark.parametrize("name", ondemand_case_names())
def test_exact_mode_keeps_certain_labels_and_stays_close_to_the_reference(name):
    # mode 2 integrates every non-certain vertex alone to a certain terminal
    # (the reference's pass-3 rule); certain regions keep their labels, and the
    # result differs from the reference pipeline only where path compression
    # labelled a vertex differently
    f, (err, gthr, iters, presimp), ref = load_ondemand(name)
    cert, seeds, _ = oracle_lib.certain_regions(f, oracle_lib.lstar_pairs(f), presimp)
    exact, _ = oracle_lib.integrate(f, cert, seeds, err, gthr, iters, mode=2)
    assert (exact[cert != -1] == cert[cert != -1]).all()
    _, lab2, _, _ = oracle_lib.ondemand_asc(f, err, gthr, iters, presimp, mode=2)
    mism = float((lab2 != ref["cleaned"]).mean())
    assert mism < 0.01, mism


def test_exact_mode_differs_from_path_compression_somewhere():
    # a planted no-op would make mode 2 equal mode 1; on the golden whose
    # streamlines run out of budget, path compression labels some vertices
    # otherwise
    f, (err, gthr, iters, presimp), _ = load_ondemand("budget_24x30x20")
    _, l1, _, _ = oracle_lib.ondemand_asc(f, err, gthr, iters, presimp, mode=1)
    _, l2, _, _ = oracle_lib.ondemand_asc(f, err, gthr, iters, presimp, mode=2)
    assert 0 < int((l1 != l2).sum())


@pytest.mark.parametrize("mode", [1, 2])
def test_two_basin_closed_form(mode):
    # f = min(|p - a|^2, |p - b|^2): the flow runs straight to the nearer
    # centre, so away from the bisector plane every vertex's FP64 label is the
    # basin of the nearer centre (the path-compressing mode 1 and the exact
    # mode 2 alike); a mistake in the step rule, the interpolation or the
    # labelling of the end vertex moves vertices across
    Z, Y, X = 24, 30, 40
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    a, b = np.array([11.3, 14.6, 10.2]), np.array([10.8, 13.9, 29.7])  # (z, y, x)
    da = (z - a[0]) ** 2 + (y - a[1]) ** 2 + (x - a[2]) ** 2
    db = (z - b[0]) ** 2 + (y - b[1]) ** 2 + (x - b[2]) ** 2
    f = np.minimum(da, db).astype(np.float32)
    _, lab, _, _ = oracle_lib.ondemand_asc(f, 1e-3, 1e-6, 4096, 0.0, mode=mode)
    lab = lab.reshape(Z, Y, X)
    la = lab[int(round(a[0])), int(round(a[1])), int(round(a[2]))]
    lb = lab[int(round(b[0])), int(round(b[1])), int(round(b[2]))]
    assert la != lb
    # signed distance to the bisector plane, in voxels
    n = b - a
    s = (np.stack([z, y, x], -1) - (a + b) / 2) @ n / np.linalg.norm(n)
    far_a, far_b = s < -2.0, s > 2.0
    assert far_a.sum() > 5000 and far_b.sum() > 5000
    assert (lab[far_a] == la).all(), int((lab[far_a] != la).sum())
    assert (lab[far_b] == lb).all(), int((lab[far_b] != lb).sum())
