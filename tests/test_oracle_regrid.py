"""CPU checks of the refinement oracle (tagging and cell_cons_interp) against
what the methods fix, and of the grid generation in libmgx.so (a host-only
entry point) against AMReX's layout guarantees.

Pins of cons_interp (AMReX is not vendored):
  * constants and linear fields are reproduced exactly (1e-14), including
    next to ext_dir faces, where the ghost holds the face value;
  * the 8 children of a coarse cell average to it (1e-14);
  * a field varying along one direction keeps every child inside the range
    of its coarse cell and the two neighbours along that direction;
  * the limiter is shared by the components: a component with an extremum
    scales the slope of a linear one (hand-computed value).
Pins of tag_cells: hand-built fields with known tagged sets.
"""
import numpy as np
import pytest

from oracle import regrid as R


def _lin(n, a, b, off=0.0):
    """u = a + b . x at the cell centres of an (n[2], n[1], n[0]) grid (h = 1)."""
    z, y, x = np.meshgrid(*[np.arange(n[d]) + 0.5 + off for d in (2, 1, 0)], indexing="ij")
    return a + b[0] * x + b[1] * y + b[2] * z


def test_cons_interp_constant_and_linear_exact():
    n = (6, 5, 7)
    c = np.full((1, n[2], n[1], n[0]), 3.25)
    f = R.cons_interp(R.ghosted(c, (1, 1, 1)), (1, 1, 1))
    assert np.all(f == 3.25)
    b = (0.3, -1.1, 0.7)
    u = _lin(n, 2.0, b)[None]
    # ghost planes: the linear extension (values the caller's physical fill
    # would hand in)
    cg = np.pad(u, ((0, 0), (1, 1), (1, 1), (1, 1)))
    zz, yy, xx = np.meshgrid(*[np.arange(-1, n[d] + 1) + 0.5 for d in (2, 1, 0)], indexing="ij")
    cg[0] = 2.0 + b[0] * xx + b[1] * yy + b[2] * zz
    f = R.cons_interp(cg, (0, 0, 0), [[2] * 6])
    fine = np.meshgrid(*[(np.arange(2 * n[d]) + 0.5) / 2 for d in (2, 1, 0)], indexing="ij")
    exact = 2.0 + b[0] * fine[2] + b[1] * fine[1] + b[2] * fine[0]
    assert np.max(np.abs(f[0] - exact)) < 1e-13
    # ext_dir: the ghost holds the value on the domain face, the one-sided
    # slope still reproduces the linear field
    cg2 = cg.copy()
    cg2[0, :, :, 0] = 2.0 + b[0] * 0.0 + b[1] * yy[:, :, 0] + b[2] * zz[:, :, 0]
    cg2[0, :, :, -1] = 2.0 + b[0] * n[0] + b[1] * yy[:, :, -1] + b[2] * zz[:, :, -1]
    f2 = R.cons_interp(cg2, (0, 0, 0), [[3, 2, 2, 3, 2, 2]])
    assert np.max(np.abs(f2[0] - exact)) < 1e-13
    # the plain central slope there would not be exact
    f3 = R.cons_interp(cg2, (0, 0, 0), [[2] * 6])
    assert np.max(np.abs(f3[0] - exact)) > 1e-3


def test_cons_interp_conserves_and_bounds():
    rng = np.random.default_rng(3)
    n = (8, 6, 4)
    c = rng.standard_normal((2, n[2], n[1], n[0]))
    f = R.cons_interp(R.ghosted(c, (1, 1, 1)), (1, 1, 1))
    avg = f.reshape(2, n[2], 2, n[1], 2, n[0], 2).mean(axis=(2, 4, 6))
    assert np.max(np.abs(avg - c)) < 1e-14
    # variation along x only: children stay within their 1-D neighbourhood
    line = rng.standard_normal(n[0])
    c1 = np.broadcast_to(line, (1, n[2], n[1], n[0])).copy()
    f1 = R.cons_interp(R.ghosted(c1, (1, 1, 1)), (1, 1, 1))[0, 0, 0]
    lo = np.minimum(np.minimum(np.roll(line, 1), line), np.roll(line, -1))
    hi = np.maximum(np.maximum(np.roll(line, 1), line), np.roll(line, -1))
    assert np.all(f1 >= np.repeat(lo, 2) - 1e-15) and np.all(f1 <= np.repeat(hi, 2) + 1e-15)


def test_cons_interp_shared_limiter():
    # comp 0 linear in x (slope 1); comp 1 has a peak at the middle cell
    n = (5, 2, 2)
    x = np.arange(n[0]) + 0.5
    c = np.zeros((2, n[2], n[1], n[0]))
    c[0] = x
    c[1] = np.array([0.0, 1.0, 1.5, 1.0, 0.0])
    cg = R.ghosted(c, (0, 1, 1))
    cg[0, :, :, 0] = -0.5
    cg[0, :, :, -1] = 5.5
    f = R.cons_interp(cg, (0, 1, 1), [[2] * 6, [2] * 6])
    # cell 1: comp1 central 0.75, forward 1.0, backward 2.0 -> factor 1;
    # cell 2: comp1 central 0, no limit; forward/backward of comp1 differ
    # in sign -> comp1 limits nothing (c = 0); cell 3 mirrors cell 1.
    # cell 1 of comp0: factor min(1, 1) -> exact
    assert abs(f[0, 0, 0, 2] - 1.25) < 1e-15 and abs(f[0, 0, 0, 3] - 1.75) < 1e-15
    # make comp1's cell 1 steep on one side: backward 2*(1.0-0.9)=0.2,
    # central (1.5-0.9)/2=0.3 -> factor 2/3 also scales comp0's slope
    c[1] = np.array([0.9, 1.0, 1.5, 1.0, 0.0])
    cg = R.ghosted(c, (0, 1, 1))
    cg[0, :, :, 0] = -0.5
    cg[0, :, :, -1] = 5.5
    f = R.cons_interp(cg, (0, 1, 1), [[2] * 6, [2] * 6])
    assert abs(f[0, 0, 0, 2] - (1.5 - 0.25 * 2.0 / 3.0)) < 1e-15
    assert abs(f[1, 0, 0, 2] - (1.0 - 0.25 * 0.2)) < 1e-15


def test_tag_cells_hand_built():
    n = (8, 8, 8)
    rho = np.ones((n[2] + 2, n[1] + 2, n[0] + 2))
    rho[4, 5, 6] = 2.0  # cell (x=5, y=4, z=3)
    t = R.tag_cells(rho, rhoerr=1.5)
    assert t.sum() == 1 and t[3, 4, 5] == 1
    t = R.tag_cells(rho, gradrhoerr=1.0)
    want = {(3, 4, 5), (2, 4, 5), (4, 4, 5), (3, 3, 5), (3, 5, 5), (3, 4, 4), (3, 4, 6)}
    assert set(map(tuple, np.argwhere(t))) == want
    assert R.tag_cells(rho, gradrhoerr=1.0 + 1e-12).sum() == 0  # >= is inclusive
    assert R.tag_cells(rho, rhoerr=2.0).sum() == 0              # > is strict
    t = R.tag_cells(rho, region=((0.25, 0.0, 0.5), (0.5, 1.0, 0.75)), dx=(1 / 8,) * 3)
    # centres (i+0.5)/8 in [0.25, 0.5] -> i = 2, 3; z in [0.5, 0.75] -> k = 4, 5
    assert set(map(tuple, np.argwhere(t))) == {(k, j, i) for k in (4, 5) for j in range(8)
                                              for i in (2, 3)}


def _sphere_tags(n, c, r):
    z, y, x = np.meshgrid(*[np.arange(n[d]) + 0.5 for d in (2, 1, 0)], indexing="ij")
    return (((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) < r * r).astype(np.uint8)


@pytest.mark.parametrize("eff", [0.5, 0.7, 0.9])
def test_cluster_tags_layout_rules(eff):
    from incflo_amd import mgx
    n = (64, 48, 32)
    t = _sphere_tags(n, (20, 24, 16), 9) | _sphere_tags(n, (50, 10, 8), 5)
    for per in ((1, 1, 1), (0, 1, 0)):
        boxes = mgx.cluster_tags(t, periodic=per, n_error_buf=1, blocking_factor=8,
                                 max_grid_size=32, grid_eff=eff)
        frac = R.check_grids(boxes, t, 1, 8, 32, per)
        assert 0.2 < frac <= 1.0
    # a higher efficiency never covers more cells
    vols = []
    for e in (0.3, 0.95):
        boxes = mgx.cluster_tags(t, grid_eff=e)
        vols.append(sum(np.prod([h[d] - l[d] + 1 for d in range(3)]) for l, h in boxes))
    assert vols[1] <= vols[0]


def test_cluster_tags_edge_cases():
    from incflo_amd import mgx
    n = (16, 16, 16)
    t = np.zeros((n[2], n[1], n[0]), np.uint8)
    assert mgx.cluster_tags(t) == []
    t[:] = 1
    boxes = mgx.cluster_tags(t, max_grid_size=16)
    R.check_grids(boxes, t, 1, 8, 16, (1, 1, 1))
    assert len(boxes) == 8  # 32^3 fine cells in 16^3 pieces
    t[:] = 0
    t[0, 0, 0] = 1  # periodic buffering wraps to the far corners
    boxes = mgx.cluster_tags(t, periodic=(1, 1, 1))
    R.check_grids(boxes, t, 1, 8, 32, (1, 1, 1))
    assert len(boxes) == 8
    boxes = mgx.cluster_tags(t, periodic=(0, 0, 0))
    assert boxes == [((0, 0, 0), (7, 7, 7))]
    with pytest.raises(mgx.MgxError):
        mgx.cluster_tags(t, blocking_factor=6, max_grid_size=32)  # 32 is no multiple of 6
    with pytest.raises(mgx.MgxError):
        mgx.cluster_tags(t, grid_eff=0.0)
    with pytest.raises(mgx.MgxError):
        mgx.cluster_tags(np.ones((16, 16, 16), np.uint8), max_grid_size=8, max_boxes=4)
