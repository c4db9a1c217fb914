"""The CPU oracle of the accurate 1-manifolds (needASC1 / needDSC1), pinned
against the reference's own runs: the restriction labels the digitized
streamlines leave and the conforming .grad, bit-exact, alone and after the
accurate 3-manifold labelling (which zeroes the shared gradient field at its
region extrema before the streamlines run)."""
import hashlib

import numpy as np
import pytest

import gradcheck
import oracle_lib
from golden import lines_case_names, load_lines


@pytest.mark.parametrize("name,flags", lines_case_names())
def test_oracle_lines_match_reference(name, flags):
    f, (err, gthr, iters, presimp), ref_r, ref_sha = load_lines(name, flags)
    Z, Y, X = f.shape
    g, lab, r, st = oracle_lib.ondemand(f, *flags, err, gthr, iters, presimp, mode=0)
    assert (r == ref_r).all(), int((r != ref_r).sum())
    assert hashlib.sha256(g.tobytes()).hexdigest() == ref_sha
    gradcheck.check_consistency(g, X, Y, Z)


def test_lines_label_only_their_cell_types():
    # ascending lines cross voxel faces: hexes and quads (and the 2-saddle
    # quads); descending lines cross dual faces: vertices and edges
    f, (err, gthr, iters, presimp), _, _ = load_lines("smooth_34x33x32", (1, 1, 0))
    _, _, r, st = oracle_lib.ondemand(f, 1, 1, 0, err, gthr, iters, presimp, mode=0)
    Z, Y, X = f.shape
    M = (2 * Z - 1, 2 * Y - 1, 2 * X - 1)
    dim = sum(np.indices(M)[a] % 2 for a in range(3)).ravel()
    assert set(np.unique(dim[r == 4])) <= {2, 3} and set(np.unique(dim[r == 5])) <= {0, 1}
    assert st[4] > 0 and st[12] > 0 and (r == 4).sum() > 0 and (r == 5).sum() > 0


def test_saddles_labelled_and_faces_next_to_their_cells():
    # every saddle takes the target after its streamlines; a face is labelled
    # only right after the cell a streamline leaves through it, so every
    # labelled face (quad ASC, edge DSC) that is not a saddle has a labelled
    # neighbour across it (a cofacet hex ASC, a facet vertex DSC)
    f, (err, gthr, iters, presimp), _, _ = load_lines("smooth_presimp_40", (1, 0, 0))
    Z, Y, X = f.shape
    base = oracle_lib.lstar_pairs(f)
    M = (2 * Z - 1, 2 * Y - 1, 2 * X - 1)
    crit = (((base >> 2) & 7) == 7).reshape(M)
    idx = np.indices(M)
    dim = sum(idx[a] % 2 for a in range(3))
    for asc, target, sdim in ((1, 4, 2), (0, 5, 1)):
        r, st = oracle_lib.saddle_lines(f, base, asc, err=err, gthr=gthr, iters=iters, mode=0)
        r = r.reshape(M)
        sad = crit & (dim == sdim)
        assert sad.sum() == st[0] > 0 and (r[sad] == target).all()
        faces = (r == target) & (dim == sdim) & ~sad
        ok = np.zeros(M, bool)
        for a in range(3):  # numpy axis a; the face crosses it where its coordinate is even (ASC) / odd (DSC)
            lo = [slice(None)] * 3
            hi = [slice(None)] * 3
            lo[a] = slice(0, -1)
            hi[a] = slice(1, None)
            nb = np.zeros(M, bool)
            nb[tuple(lo)] |= r[tuple(hi)] == target
            nb[tuple(hi)] |= r[tuple(lo)] == target
            ok |= nb & (idx[a] % 2 == (0 if asc else 1))
        assert faces.sum() > 0 and ok[faces].all(), int((~ok[faces]).sum())
