"""Region boundary kinds (ComputeRegionBoundaryKind, the `_edges_` file of
Accurate3DMorseLabeling), the oracle pinned by closed forms.

The reference's own code cannot be built: it assigns through DenseLabeling's
const operator[] (gi_labeling_to_bounary_labeling.h:287 ff., gi_labeling.h:
170-171), so no run of it exists to compare with.  Labellings whose kinds
follow from geometry pin every rule instead: one region (all 0), two
half-spaces (a wall of 1), four quadrants (a line of 2), eight octants (a
single 3), and unlabelled vertices (0 on their edges)."""
import numpy as np

import oracle_lib

SHAPE = (6, 7, 8)  # Z, Y, X


def _kinds(lab):
    Z, Y, X = lab.shape
    return oracle_lib.region_boundary_kind(lab, lab.shape).reshape(2 * Z - 1, 2 * Y - 1, 2 * X - 1)


def _cell_coords(Z, Y, X):
    return np.indices((2 * Z - 1, 2 * Y - 1, 2 * X - 1))


def test_one_region_is_all_zero():
    assert (_kinds(np.full(SHAPE, 5, np.int32)) == 0).all()


def test_two_half_spaces_give_a_wall_of_ones():
    Z, Y, X = SHAPE
    k = 3
    lab = np.where(np.arange(X)[None, None, :] < k, 1, 2).astype(np.int32) * np.ones(SHAPE, np.int32)
    kd = _kinds(lab)
    cz, cy, cx = _cell_coords(Z, Y, X)
    wall = cx == 2 * k - 1  # every cell of the complex between x = k-1 and x = k
    assert (kd[wall] == 1).all() and (kd[~wall] == 0).all()


def test_four_quadrants_give_a_line_of_twos():
    Z, Y, X = SHAPE
    k, j = 3, 4
    x = np.arange(X)[None, None, :]
    y = np.arange(Y)[None, :, None]
    lab = ((x < k) * 1 + (y < j) * 2).astype(np.int32) * np.ones(SHAPE, np.int32)
    kd = _kinds(lab)
    cz, cy, cx = _cell_coords(Z, Y, X)
    line = (cx == 2 * k - 1) & (cy == 2 * j - 1)  # quads and hexes around the junction line
    walls = ((cx == 2 * k - 1) | (cy == 2 * j - 1)) & ~line
    assert (kd[line & (cz % 2 == 0)] == 2).all()  # quads with four labels
    assert (kd[line & (cz % 2 == 1)] == 2).all()  # hexes: two quads of kind 2, so the largest kind
    assert (kd[walls] == 1).all() and (kd[~(line | walls)] == 0).all()


def test_eight_octants_meet_in_one_hex_of_three():
    Z, Y, X = SHAPE
    k, j, l = 3, 4, 2
    x = np.arange(X)[None, None, :]
    y = np.arange(Y)[None, :, None]
    z = np.arange(Z)[:, None, None]
    lab = ((x < k) * 1 + (y < j) * 2 + (z < l) * 4).astype(np.int32)
    kd = _kinds(lab)
    cz, cy, cx = _cell_coords(Z, Y, X)
    centre = (cx == 2 * k - 1) & (cy == 2 * j - 1) & (cz == 2 * l - 1)
    assert kd[centre].tolist() == [3]
    assert (kd[~centre] <= 2).all()
    # its six faces each see four regions
    for dz, dy, dx in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        assert kd[2 * l - 1 + dz, 2 * j - 1 + dy, 2 * k - 1 + dx] == 2


def test_unlabelled_vertices_leave_their_edges_zero():
    lab = np.ones(SHAPE, np.int32)
    lab[:, :, :3] = 2
    lab[2, 3, 3] = -1  # on the wall side x = 3
    kd = _kinds(lab)
    # the wall edge from (x=2) to (x=3) at (y=3, z=2) has an unlabelled end
    assert kd[4, 6, 5] == 0
    assert kd[4, 4, 5] == 1
