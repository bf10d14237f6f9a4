"""GPU parity of dynamic refinement (mgx_tag_cells, mgx_fill_fine_level and
the regrid driver) against the oracle (oracle/regrid.py), through the C ABI.

Tolerances: tags bit-exact (integer work); the fine-level fill within 1e-14
of the oracle's cell_cons_interp (the same FP64 arithmetic in another order)
and bit-exact where the old fine level supplies the values (copies).
"""
import sys

import numpy as np
import pytest

from incflo_amd import boxes as B
from oracle import regrid as R

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from incflo_amd import mgx
    c = mgx.Context(0)
    yield c
    c.close()


def _boxed_ghosted(cg, boxes, ng, ncomp):
    """boxed field whose every cell (valid and ghost) comes from the ghosted
    canonical array cg (ncomp, nz+2, ny+2, nx+2)"""
    return B.set_ghosts_from(np.zeros(B.field_size(boxes, ng, ncomp)), boxes, ng, cg,
                             ncomp=ncomp, gw=1)


def _mask(boxes, shape):
    m = np.zeros(shape, dtype=bool)
    for lo, hi in boxes:
        m[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
    return m


@pytest.mark.parametrize("per", [(1, 1, 1), (0, 1, 0)])
def test_tags_match_oracle(ctx, per):
    from incflo_amd import mgx
    n = (24, 16, 20)
    _say(f"case tags periodic={per}")
    bxs = B.chop_domain(n, (8, 16, 7))
    lev = mgx.make_level(n, bxs, periodic=per, prob_hi=(1.5, 1.0, 1.25))
    rng = np.random.default_rng(4)
    rho = (1.0 + rng.random((1, n[2], n[1], n[0])))
    rho[0, 5:9, 4:10, 6:12] += 1.5
    # periodic images, and the caller's boundary values along the bounded directions
    cg = R.ghosted(rho, per)
    if not per[0]:
        cg[0, :, :, 0] = 1.1
        cg[0, :, :, -1] = 1.3
    if not per[2]:
        cg[0, 0] = 0.9
        cg[0, -1] = 1.2
    rd = ctx.array(_boxed_ghosted(cg, bxs, 1, 1))
    region = ((0.1, 0.2, 0.3), (0.4, 0.5, 0.6))
    for kw in (dict(rhoerr=2.0), dict(gradrhoerr=0.9), dict(region=region),
               dict(rhoerr=2.2, gradrhoerr=1.2, region=region)):
        t = mgx.tag_cells(ctx, lev, rd, 1, **kw)
        o = R.tag_cells(cg[0], kw.get("rhoerr"), kw.get("gradrhoerr"), kw.get("region"),
                        dx=(1.5 / n[0], 1.0 / n[1], 1.25 / n[2]))
        assert t.sum() > 0 and np.array_equal(t, o), kw


def _fill_case(ctx, per, bcrec, old=False, seed=2):
    from incflo_amd import mgx
    nc = (16, 12, 8)
    ncomp = 3
    rng = np.random.default_rng(seed)
    zz, yy, xx = np.meshgrid(*[(np.arange(nc[d]) + 0.5) / nc[d] for d in (2, 1, 0)],
                             indexing="ij")
    c = np.stack([np.sin(2 * np.pi * xx) + 0.3 * rng.standard_normal(xx.shape),
                  np.cos(2 * np.pi * yy) * zz, 0.5 * rng.standard_normal(xx.shape)])
    cg = R.ghosted(c, per)
    for d in range(3):
        if not per[d]:  # caller's boundary values in the ghost planes
            ax = 3 - d
            lo = [slice(None)] * 4
            hi = [slice(None)] * 4
            lo[ax], hi[ax] = 0, -1
            cg[tuple(lo)] = 0.7
            cg[tuple(hi)] = -0.4
    cb = B.chop_domain(nc, 8)
    lev_c = mgx.make_level(nc, cb, periodic=per)
    cd = ctx.array(_boxed_ghosted(cg, cb, 1, ncomp))
    fine_boxes = [((0, 0, 0), (7, 15, 7)), ((8, 0, 0), (23, 7, 15)), ((24, 16, 8), (31, 23, 15))]
    lev_f = mgx.make_level([2 * v for v in nc], fine_boxes, periodic=per)
    fd = ctx.array(B.field_size(fine_boxes, 1, ncomp))
    expect = R.cons_interp(cg, per, bcrec)
    fshape = (2 * nc[2], 2 * nc[1], 2 * nc[0])
    kw = {}
    own = np.zeros(fshape, dtype=bool)
    if old:
        old_boxes = [((0, 0, 0), (15, 7, 7)), ((16, 8, 8), (31, 23, 15))]
        of = rng.standard_normal((ncomp,) + fshape)
        lev_o = mgx.make_level([2 * v for v in nc], old_boxes, periodic=per)
        od = ctx.array(B.to_boxed(of, old_boxes, 2, periodic=per))
        kw = dict(old_fine=lev_o, old_field=od, of_ng=2)
        own = _mask(old_boxes, fshape)
        expect = np.where(own[None], of, expect)
    mgx.fill_fine_level(ctx, lev_c, cd, 1, ncomp, lev_f, fd, 1, bcrec=bcrec, **kw)
    got = B.from_boxed(fd.download(), fine_boxes, 1, fshape, ncomp=ncomp)
    m = _mask(fine_boxes, fshape)
    return got, expect, m, own


@pytest.mark.parametrize("per", [(1, 1, 1), (0, 1, 0)])
def test_fill_from_coarse_matches_oracle(ctx, per):
    _say(f"case fill from coarse periodic={per}")
    bc = [[3 if not per[d % 3] else 0 for d in range(6)], [2 if not per[d % 3] else 0 for d in range(6)],
          [4 if not per[d % 3] else 0 for d in range(6)]]
    got, expect, m, _ = _fill_case(ctx, per, bc)
    err = np.abs(got - expect)[:, m].max() / np.abs(expect[:, m]).max()
    assert err < 1e-14
    # conservation: the 8 children of every coarse cell under the fine boxes
    # (aligned to 2) average to it
    nz, ny, nx = m.shape
    avg = got.reshape(3, nz // 2, 2, ny // 2, 2, nx // 2, 2).mean(axis=(2, 4, 6))
    mc = m[::2, ::2, ::2]
    cc = expect.reshape(3, nz // 2, 2, ny // 2, 2, nx // 2, 2).mean(axis=(2, 4, 6))
    assert np.abs(avg - cc)[:, mc].max() < 1e-14


def test_fill_keeps_old_fine_values(ctx):
    _say("case fill from old fine + coarse")
    got, expect, m, own = _fill_case(ctx, (1, 1, 1), None, old=True)
    both = m & own
    assert both.sum() > 0 and (m & ~own).sum() > 0
    assert np.array_equal(got[:, both], expect[:, both])  # copies are exact
    rest = m & ~own
    assert np.abs(got - expect)[:, rest].max() < 1e-14 * np.abs(expect[:, rest]).max()


def test_regrid_follows_a_moving_bubble_and_projects(ctx):
    """Two regrids of a density bubble that moves; the second keeps the old
    fine data where the levels overlap.  The composite projection then runs on
    the new hierarchy and converges."""
    from incflo_amd import mgx
    from incflo_amd import regrid as G
    nc = (32, 32, 32)
    _say("case regrid driver 32^3 -> fine")
    cb = B.chop_domain(nc, 16)
    lev_c = mgx.make_level(nc, cb)
    zz, yy, xx = np.meshgrid(*[(np.arange(nc[d]) + 0.5) / nc[d] for d in (2, 1, 0)],
                             indexing="ij")

    def bubble(cx):
        return 1.0 + 2.0 * (((xx - cx) ** 2 + (yy - 0.5) ** 2 + (zz - 0.5) ** 2) < 0.15 ** 2)

    par = G.RegridParams(rhoerr=1.5, n_error_buf=1, blocking_factor=8, max_grid_size=16,
                         grid_eff=0.7)
    rho0 = bubble(0.35)[None]
    vel = np.stack([np.sin(2 * np.pi * yy), np.sin(2 * np.pi * zz), np.sin(2 * np.pi * xx)])
    rd = ctx.array(_boxed_ghosted(R.ghosted(rho0, (1, 1, 1)), cb, 1, 1))
    vd = ctx.array(_boxed_ghosted(R.ghosted(vel, (1, 1, 1)), cb, 1, 3))
    fine1, f1 = G.regrid(ctx, lev_c, rd, 1, par, {"density": (rd, 1, 1, None),
                                                  "velocity": (vd, 1, 3, None)})
    b1 = G.boxes_of(fine1)
    R.check_grids(b1, R.tag_cells(R.ghosted(rho0, (1, 1, 1))[0], 1.5), 1, 8, 16, (1, 1, 1))
    # the bubble moves; regrid again, keeping the old fine data
    rho1 = bubble(0.6)[None]
    rd2 = ctx.array(_boxed_ghosted(R.ghosted(rho1, (1, 1, 1)), cb, 1, 1))
    fine2, f2 = G.regrid(ctx, lev_c, rd2, 1, par,
                         {"density": (rd2, 1, 1, None), "velocity": (vd, 1, 3, None)},
                         old_fine=fine1, old_fields={"density": (f1["density"], 1),
                                                     "velocity": (f1["velocity"], 1)})
    b2 = G.boxes_of(fine2)
    assert b2 != b1
    R.check_grids(b2, R.tag_cells(R.ghosted(rho1, (1, 1, 1))[0], 1.5), 1, 8, 16, (1, 1, 1))
    fshape = (64, 64, 64)
    d1 = B.from_boxed(f1["density"].download(), b1, 1, fshape)
    d2 = B.from_boxed(f2["density"].download(), b2, 1, fshape)
    m1, m2 = _mask(b1, fshape), _mask(b2, fshape)
    assert (m1 & m2).sum() > 0
    assert np.array_equal(d2[m1 & m2], d1[m1 & m2])
    new = m2 & ~m1
    exp = R.cons_interp(R.ghosted(rho1, (1, 1, 1)), (1, 1, 1))[0]
    assert np.abs(d2[new] - exp[new]).max() < 1e-14 * np.abs(exp).max()
    # the two-level projection on the regridded hierarchy, with the coarse
    # density under the fine level averaged down from it (as incflo keeps it)
    mgx.amr_average_down(ctx, lev_c, fine2, f2["density"], 1, rd2, 1)
    rc = B.from_boxed(rd2.download(), cb, 1, nc[::-1])
    P = mgx.AmrNodalProjector(ctx, [lev_c, fine2])
    sc = ctx.array(B.to_boxed(0.01 / rc, cb, 0))
    rf = B.from_boxed(f2["density"].download(), b2, 1, fshape)
    sf = ctx.array(B.to_boxed(np.where(m2, 0.01 / np.where(m2, rf, 1.0), 0.0), b2, 0))
    P.set_sigma([sc, sf], 0)
    vf = B.from_boxed(f2["velocity"].download(), b2, 1, fshape, ncomp=3)
    vcd = ctx.array(B.to_boxed(vel, cb, 1))
    vfd = ctx.array(B.to_boxed(vf, b2, 1))
    info = P.project([vcd, vfd], 1, 1e-10, 0.0, mode=mgx.MODE_FP64)
    assert info.converged and info.rel_resid <= 1e-10 and info.iters <= 20
    P.close()


def test_three_level_regrid_nests_and_projects(ctx):
    """Two regrids of every level (regrid_levels, max_level 2) over a density
    bubble: the first makes level 1, the second level 2 from level 1's tags
    and level 1 again around it.  The hierarchy nests as the composite
    projection requires (the oracle's construction accepts it), each level's
    boxes cover the buffered tags of the level below, level 2 holds the
    conservative interpolation of the new level 1, and the three-level
    composite projection converges on it."""
    import oracle
    from incflo_amd import mgx
    from incflo_amd import regrid as G
    from oracle.amr import AmrLevels
    nc = (32, 32, 32)
    per = (1, 1, 1)
    _say("case regrid three levels 32^3")
    cb = B.chop_domain(nc, 16)
    lev0 = mgx.make_level(nc, cb)
    zz, yy, xx = np.meshgrid(*[(np.arange(nc[d]) + 0.5) / nc[d] for d in (2, 1, 0)], indexing="ij")
    rho0 = (1.0 + 2.0 * (((xx - 0.45) ** 2 + (yy - 0.5) ** 2 + (zz - 0.5) ** 2) < 0.2 ** 2))[None]
    vel = np.stack([np.sin(2 * np.pi * yy), np.sin(2 * np.pi * zz), np.sin(2 * np.pi * xx)])
    rd = ctx.array(_boxed_ghosted(R.ghosted(rho0, per), cb, 1, 1))
    vd = ctx.array(_boxed_ghosted(R.ghosted(vel, per), cb, 1, 3))
    pars = [G.RegridParams(rhoerr=1.5, n_error_buf=1, blocking_factor=8, max_grid_size=16),
            G.RegridParams(rhoerr=2.5, n_error_buf=1, blocking_factor=8, max_grid_size=16)]
    f0 = {"density": (rd, 1, 1, None), "velocity": (vd, 1, 3, None)}
    levs, flds = G.regrid_levels(ctx, [lev0], [(rd, 1)], pars, [f0], max_level=2)
    assert len(levs) == 2
    levs2, flds2 = G.regrid_levels(ctx, levs, [(rd, 1), (flds[1]["density"][0], 1)], pars, flds,
                                   max_level=2)
    assert len(levs2) == 3
    b1, b2 = G.boxes_of(levs2[1]), G.boxes_of(levs2[2])
    _say(f"  level 1: {len(b1)} boxes, level 2: {len(b2)} boxes")
    gc = oracle.Geom.make(nc)
    A = AmrLevels(gc, [[lo + hi for lo, hi in b1], [lo + hi for lo, hi in b2]])
    t0 = mgx.tag_cells(ctx, lev0, rd, 1, 1.5)
    R.check_grids(b1, t0, 1, 8, 16, per)
    t1 = mgx.tag_cells(ctx, levs[1], flds[1]["density"][0], 1, 2.5)
    assert t1.sum() > 0
    R.check_grids(b2, t1, 1, 8, 16, per)
    # level 2 from the new level 1 (no old level 2)
    s1, s2 = (64, 64, 64), (128, 128, 128)
    d1 = B.from_boxed(flds2[1]["density"][0].download(), b1, 1, s1)
    d2 = B.from_boxed(flds2[2]["density"][0].download(), b2, 1, s2)
    exp2 = R.cons_interp(R.ghosted(d1[None], per), per)[0]
    m2 = _mask(b2, s2)
    assert np.abs(d2[m2] - exp2[m2]).max() < 1e-14 * np.abs(exp2[m2]).max()
    # the three-level projection, the density averaged down level by level
    mgx.amr_average_down(ctx, levs2[1], levs2[2], flds2[2]["density"][0], 1,
                         flds2[1]["density"][0], 1)
    mgx.amr_average_down(ctx, lev0, levs2[1], flds2[1]["density"][0], 1, rd, 1)
    rc = B.from_boxed(rd.download(), cb, 1, nc[::-1])
    d1 = B.from_boxed(flds2[1]["density"][0].download(), b1, 1, s1)
    m1 = _mask(b1, s1)
    sig = [ctx.array(B.to_boxed(0.01 / rc, cb, 0)),
           ctx.array(B.to_boxed(np.where(m1, 0.01 / np.where(m1, d1, 1.0), 0.0), b1, 0)),
           ctx.array(B.to_boxed(np.where(m2, 0.01 / np.where(m2, d2, 1.0), 0.0), b2, 0))]
    P = mgx.AmrNodalProjector(ctx, levs2)
    P.set_sigma(sig, 0)
    v1 = B.from_boxed(flds2[1]["velocity"][0].download(), b1, 1, s1, ncomp=3)
    v2 = B.from_boxed(flds2[2]["velocity"][0].download(), b2, 1, s2, ncomp=3)
    vels = [ctx.array(B.to_boxed(vel, cb, 1)), ctx.array(B.to_boxed(v1, b1, 1)),
            ctx.array(B.to_boxed(v2, b2, 1))]
    info = P.project(vels, 1, 1e-10, 0.0, mode=mgx.MODE_FP64)
    _say(f"  three-level projection: {info.iters} composite V-cycles")
    assert info.converged and info.rel_resid <= 1e-10 and info.iters <= 25
    P.close()
    assert A.dof[2].sum() > 0
