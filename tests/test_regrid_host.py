"""Host logic of the multi-level regrid driver (incflo_amd/regrid.py): the
projection of the new finer grids onto the tags of the level two below
(AmrMesh::MakeNewGrids: coarsen to the level between, grow by n_proper there,
coarsen again).  Expected tags written out by hand for each case."""
import numpy as np

from incflo_amd.regrid import project_finer


def _box_tags(n, lo, hi):
    t = np.zeros((n[2], n[1], n[0]), bool)
    t[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
    return t


def test_interior_box():
    # level l+2 cells 8..15 -> level l+1 4..7 -> grown 3..8 -> level l 1..4
    t = project_finer([((8, 8, 8), (15, 15, 15))], (8, 8, 8), (1, 1, 1))
    assert np.array_equal(t, _box_tags((8, 8, 8), (1, 1, 1), (4, 4, 4)))


def test_periodic_wrap_and_bounded_clip():
    # x: 0..7 -> 0..3 -> -1..4 -> level l -1..2: periodic {7, 0, 1, 2}
    # y (bounded): the same box stops at 0
    t = project_finer([((0, 0, 16), (7, 7, 23))], (8, 8, 8), (1, 0, 1))
    ref = np.zeros((8, 8, 8), bool)
    for i in (7, 0, 1, 2):
        ref[3:6 + 1, 0:2 + 1, i] = True  # z: 16..23 -> 8..11 -> 7..12 -> 3..6
    assert np.array_equal(t, ref)


def test_union_of_boxes():
    b = [((8, 8, 8), (15, 15, 15)), ((20, 4, 8), (27, 11, 15))]
    t = project_finer(b, (8, 8, 8), (1, 1, 1))
    # second box: x 20..27 -> 10..13 -> 9..14 -> 4..7; y 4..11 -> 2..5 -> 1..6 -> 0..3
    ref = _box_tags((8, 8, 8), (1, 1, 1), (4, 4, 4)) | _box_tags((8, 8, 8), (4, 0, 1), (7, 3, 4))
    assert np.array_equal(t, ref)
