"""The closed form of the ldir marking that the GPU passes compute, checked on
the CPU against the reference's .grad goldens.

setAscendingManifoldDimensions (gi_topological_gradient_using_algorithms.h:395-478)
marks cells while it walks; the GPU instead records which items each pass
reaches (edges in pass 1, quads in pass 2) and derives every cell's ldir from
that afterwards (msceer_amd/csrc/ldir.hip):
  edge: 2 if reached in pass 1, else 3;
  quad: 1 if reached in pass 2; else 2 if assigned, not critical and a facet
        edge other than its pair was reached in pass 1; else 3;
  hex:  0 if critical; else 1 if a facet quad other than its pair was reached
        in pass 2; else 2 if its pair is a quad marked in pass 1; else 3.
Here that derivation runs in plain Python (breadth-first reachability over the
golden .grad) and must give the reference's ldir bits exactly.
"""
from collections import deque

import numpy as np
import pytest

from golden import load_lstar

OFF = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]


def closed_form_ldir(g, X, Y, Z):
    Mx, My, Mz = 2 * X - 1, 2 * Y - 1, 2 * Z - 1
    G = g.reshape(Mz, My, Mx)
    assigned = (G & 1) != 0
    d = (G >> 2) & 7

    def pair(c):
        dd = d[c]
        if dd == 7:
            return None
        ox, oy, oz = OFF[dd]
        return (c[0] + oz, c[1] + oy, c[2] + ox)

    def dim(c):
        return (c[0] & 1) + (c[1] & 1) + (c[2] & 1)

    def cofacets(c):
        z, y, x = c
        out = []
        for ax, (n, M) in enumerate(((x, Mx), (y, My), (z, Mz))):
            if n & 1:
                continue
            for s in (1, -1):
                if 0 <= n + s < M:
                    cc = [x, y, z]
                    cc[ax] = n + s
                    out.append((cc[2], cc[1], cc[0]))
        return out

    def facets(c):
        z, y, x = c
        out = []
        for ax, n in enumerate((x, y, z)):
            if n & 1:
                for s in (1, -1):
                    cc = [x, y, z]
                    cc[ax] = n + s
                    out.append((cc[2], cc[1], cc[0]))
        return out

    reached = [set(), set()]
    crit = [tuple(c) for c in np.argwhere((d == 7) & assigned)]
    for p, item_dim in ((0, 1), (1, 2)):
        q = deque(c for c in crit if dim(c) == item_dim)
        reached[p].update(q)
        while q:
            c = q.popleft()
            for t in cofacets(c):
                if d[t] == 7 or not assigned[t]:
                    continue
                pt = pair(t)
                if pt == c or dim(pt) != item_dim or pt in reached[p]:
                    continue
                reached[p].add(pt)
                q.append(pt)

    def marked1(t):
        if not assigned[t] or d[t] == 7:
            return False
        pt = pair(t)
        return any(e in reached[0] and e != pt for e in facets(t))

    ldir = np.full(G.shape, 3, np.uint8)
    for c in np.ndindex(G.shape):
        k = dim(c)
        if k == 1:
            ldir[c] = 2 if c in reached[0] else 3
        elif k == 2:
            ldir[c] = 1 if c in reached[1] else (2 if marked1(c) else 3)
        elif k == 3:
            if d[c] == 7:
                ldir[c] = 0
                continue
            ph = pair(c)
            if assigned[c] and any(q in reached[1] and q != ph for q in facets(c)):
                ldir[c] = 1
            elif ph is not None and pair(ph) == c and marked1(ph):
                ldir[c] = 2
    return ldir.ravel()


@pytest.mark.parametrize("name", ["noise_17x9x5", "ties_10x11x12", "plateau_33x29x21", "slab_64x64x2",
                                  "signed_zeros_9x8x7", "ramp_13x6x5", "tiny_5x2x3", "const_8x7x6"])
def test_closed_form_matches_reference_ldir(name):
    f, ref = load_lstar(name)
    Z, Y, X = f.shape
    got = closed_form_ldir(ref, X, Y, Z)
    assert (got == (ref >> 5)).all(), int((got != (ref >> 5)).sum())


@pytest.mark.parametrize("name", ["plateau_33x29x21", "noise_17x9x5"])
def test_goldens_exercise_every_branch(name):
    # each case of the closed form occurs in the reference's output, so a
    # mistake in any one of them fails the comparison above
    f, ref = load_lstar(name)
    Z, Y, X = f.shape
    Mx, My, Mz = 2 * X - 1, 2 * Y - 1, 2 * Z - 1
    zz, yy, xx = np.meshgrid(np.arange(Mz), np.arange(My), np.arange(Mx), indexing="ij")
    dim = ((zz & 1) + (yy & 1) + (xx & 1)).ravel()
    ldir = ref >> 5
    for k, values in ((1, (2, 3)), (2, (1, 2, 3)), (3, (0, 1, 2, 3))):
        for v in values:
            assert int(((dim == k) & (ldir == v)).sum()) > 0, (k, v)
