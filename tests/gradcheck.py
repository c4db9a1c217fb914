"""Independent checks of a .grad byte array (numpy): invariants the discrete
gradient must satisfy whatever implementation produced it."""
import numpy as np


def cell_dims(X, Y, Z):
    Mx, My, Mz = 2 * X - 1, 2 * Y - 1, 2 * Z - 1
    cz, cy, cx = np.meshgrid(np.arange(Mz), np.arange(My), np.arange(Mx), indexing="ij")
    return ((cx & 1) + (cy & 1) + (cz & 1)).ravel(), (cx, cy, cz)


def pair_targets(grad, X, Y, Z):
    Mx, My = 2 * X - 1, 2 * Y - 1
    code = (grad >> 2) & 7
    off = np.array([1, -1, Mx, -Mx, Mx * My, -Mx * My, 0, 0], np.int64)
    return np.arange(grad.size, dtype=np.int64) + off[code], code


def check_consistency(grad, X, Y, Z):
    """CheckGradientConsistency (gi_topological_gradient_using_algorithms.h:111-138)
    plus the Euler characteristic of the critical cells of a box (= 1)."""
    grad = np.asarray(grad).ravel()
    dims, _ = cell_dims(X, Y, Z)
    assert (grad & 1).all(), "unassigned cells"
    tgt, code = pair_targets(grad, X, Y, Z)
    crit = code == 7
    assert not (code == 6).any()
    nc = ~crit
    assert (tgt[nc] >= 0).all() and (tgt[nc] < grad.size).all()
    assert (code[tgt[nc]] != 7).all(), "paired with a critical cell"
    back, _ = pair_targets(grad, X, Y, Z)
    assert (back[tgt[nc]] == np.arange(grad.size)[nc]).all(), "pair of pair is not self"
    assert (np.abs(dims[tgt[nc]].astype(int) - dims[nc].astype(int)) == 1).all()
    counts = [int((crit & (dims == k)).sum()) for k in range(4)]
    assert counts[0] - counts[1] + counts[2] - counts[3] == 1, counts
    return counts


def ldir_reachability(grad, X, Y, Z):
    """ldir bits from their closed form: 0 for critical 3-cells; else 1 if the
    cell is reached from a critical 2-cell, else 2 if reached from a critical
    1-cell, else 3 (reachability along cofacet -> pair steps)."""
    grad = np.asarray(grad).ravel()
    Mx, My, Mz = 2 * X - 1, 2 * Y - 1, 2 * Z - 1
    dims, (cx, cy, cz) = cell_dims(X, Y, Z)
    cx, cy, cz = cx.ravel(), cy.ravel(), cz.ravel()
    tgt, code = pair_targets(grad, X, Y, Z)
    crit = code == 7
    steps = [(cx, Mx, 1), (cy, My, Mx), (cz, Mz, Mx * My)]

    def walk(start_dim):
        item = np.zeros(grad.size, bool)
        mark = np.zeros(grad.size, bool)
        frontier = np.nonzero(crit & (dims == start_dim))[0]
        item[frontier] = True
        while frontier.size:
            nxt = []
            for coord, M, st in steps:
                for sgn in (1, -1):
                    f = frontier[(coord[frontier] % 2 == 0) &
                                 ((coord[frontier] < M - 1) if sgn > 0 else (coord[frontier] > 0))]
                    t = f + sgn * st
                    ok = ~crit[t] & ((grad[t] & 1) == 1) & (tgt[t] != f)
                    t, f = t[ok], f[ok]
                    mark[t] = True
                    p = tgt[t]
                    nxt.append(p[~item[p]])
            frontier = np.unique(np.concatenate(nxt)) if nxt else np.zeros(0, np.int64)
            item[frontier] = True
        return item, mark

    item1, mark1 = walk(1)
    item2, mark2 = walk(2)
    ld = np.full(grad.size, 3, np.uint8)
    ld[(item1 | mark1)] = 2
    ld[(item2 | mark2)] = 1
    ld[crit & (dims == 3)] = 0
    return ld
