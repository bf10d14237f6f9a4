"""GPU parity of the two-level composite MAC projection (mgx_amr_mac_*)
against the oracle (oracle/mac_amr.py, a sparse direct solve of the same
composite system), through the C ABI.

The GPU solves to 1e-12 relative residual; the projected face velocities of
both levels (coarse faces under the fine level averaged down) and phi are
compared at 1e-8 of their largest value, and the composite divergence of the
GPU result, evaluated by the oracle, must be at the solver tolerance.  The
mixed FP32/FP64 mode reaches the same FP64 tolerance."""
import sys

import numpy as np
import pytest

from incflo_amd import boxes as B
from oracle import mac_amr as M

pytestmark = pytest.mark.gpu

BCD, BCN, BCR = 1, 2, 4


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from incflo_amd import mgx
    c = mgx.Context(0)
    yield c
    c.close()


def _sync(A, u, n):
    return A._sync_periodic(u, n)


def _fine_mask(nf, fboxes):
    F = np.zeros((nf[2], nf[1], nf[0]), bool)
    for lo, hi in fboxes:
        F[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
    return F


def _run(orc, ctx, nc, cboxes, fboxes, periodic, bclo=(0, 0, 0), bchi=(0, 0, 0), mode=0, seed=1):
    from incflo_amd import mgx
    ph = (1.0, nc[1] / nc[0], nc[2] / nc[0])
    gc = orc.Geom.make(nc, periodic=periodic, bclo=bclo, bchi=bchi, prob_hi=ph)
    nf = tuple(2 * v for v in nc)
    gf = orc.Geom.make(nf, periodic=periodic, bclo=bclo, bchi=bchi, prob_hi=ph)
    F = _fine_mask(nf, fboxes)
    A = M.MacAmr(gc, F)
    rng = np.random.default_rng(seed)
    bc_ = _sync(A, [0.5 + rng.random(orc.face_shape(gc, d)) for d in range(3)], A.nc)
    bf = _sync(A, [0.5 + rng.random(orc.face_shape(gf, d)) for d in range(3)], A.nf)
    uc = _sync(A, [rng.standard_normal(orc.face_shape(gc, d)) for d in range(3)], A.nc)
    uf = _sync(A, [rng.standard_normal(orc.face_shape(gf, d)) for d in range(3)], A.nf)
    ucn, ufn, phif, phic, x, div0 = A.project(bc_, bf, uc, uf)

    _say(f"amr mac: n={nc} periodic={periodic} fine boxes={len(fboxes)} mode={mode} seed={seed}")
    lev_c = mgx.make_level(nc, cboxes, periodic=periodic, prob_hi=ph)
    lev_f = mgx.make_level(nf, fboxes, periodic=periodic, prob_hi=ph)
    P = mgx.AmrMacProjector(ctx, [lev_c, lev_f])
    if not all(periodic):
        P.setDomainBC(bclo, bchi)
    P.initProjector([[ctx.array(B.faces_to_boxed(bc_[d], cboxes, 0, d)) for d in range(3)],
                     [ctx.array(B.faces_to_boxed(bf[d], fboxes, 0, d)) for d in range(3)]], 0)
    ud = [[ctx.array(B.faces_to_boxed(uc[d], cboxes, 0, d)) for d in range(3)],
          [ctx.array(B.faces_to_boxed(uf[d], fboxes, 0, d)) for d in range(3)]]
    pc = ctx.array(B.field_size(cboxes, 0, 1))
    pf = ctx.array(B.field_size(fboxes, 0, 1))
    info = P.project(ud, 0, 1e-12, 0.0, phis=[pc, pf], mode=mode)
    _say(f"  iters {info.iters} rel {info.rel_resid:.3e}")
    assert info.converged
    guc = [B.faces_from_boxed(ud[0][d].download(), cboxes, 0, d, orc.face_shape(gc, d)) for d in range(3)]
    guf = [B.faces_from_boxed(ud[1][d].download(), fboxes, 0, d, orc.face_shape(gf, d)) for d in range(3)]
    held = [B.faces_from_boxed(np.ones(B.face_field_size(fboxes, 0, d)), fboxes, 0, d,
                               orc.face_shape(gf, d)) > 0 for d in range(3)]
    gpc = B.from_boxed(pc.download(), cboxes, 0, gc.cell_shape())
    gpf = B.from_boxed(pf.download(), fboxes, 0, gf.cell_shape())
    P.close()
    return dict(A=A, gc=gc, gf=gf, F=F, ucn=ucn, ufn=ufn, phif=phif, phic=phic, div0=div0,
                guc=guc, guf=guf, held=held, gpc=gpc, gpf=gpf, info=info, bc=bc_, bf=bf)


def _check(r, tol=1e-8):
    A = r["A"]
    for d in range(3):
        sc = np.max(np.abs(r["ucn"][d]))
        err = np.max(np.abs(r["guc"][d] - r["ucn"][d]))
        assert err <= tol * sc, ("coarse", d, err, sc)
        h = r["held"][d]
        sf = np.max(np.abs(r["ufn"][d][h]))
        errf = np.max(np.abs(r["guf"][d][h] - r["ufn"][d][h]))
        assert errf <= tol * sf, ("fine", d, errf, sf)
    sp = max(np.max(np.abs(r["phic"])), np.max(np.abs(r["phif"][r["F"]])))
    assert np.max(np.abs(r["gpc"] - r["phic"])) <= tol * sp
    assert np.max(np.abs(r["gpf"][r["F"]] - r["phif"][r["F"]])) <= tol * sp
    # the GPU's projected velocity is composite divergence free (oracle's
    # operator); a periodic seam face held by one fine box only gives its
    # value to the other entry
    guf = [np.where(r["held"][d], r["guf"][d], 0.0) for d in range(3)]
    for d in range(3):
        if not A.per[d]:
            continue
        ax = 2 - d
        f0 = [slice(None)] * 3
        fn = [slice(None)] * 3
        f0[ax] = 0
        fn[ax] = -1
        h0 = r["held"][d][tuple(f0)]
        v = np.where(h0, guf[d][tuple(f0)], guf[d][tuple(fn)])
        guf[d][tuple(f0)] = v
        guf[d][tuple(fn)] = v
    div = A.divergence(guf, r["guc"])
    assert np.max(np.abs(div)) <= 1e-9 * np.max(np.abs(r["div0"]))


def test_amr_mac_periodic(orc, ctx):
    nc = (16, 12, 12)
    cboxes = B.chop_domain(nc, 8)
    fboxes = [((8, 4, 4), (15, 11, 11)), ((16, 4, 4), (23, 9, 7)), ((24, 16, 14), (31, 23, 23))]
    r = _run(orc, ctx, nc, cboxes, fboxes, (1, 1, 1))
    _check(r)


def test_amr_mac_walls_and_periodic_seam(orc, ctx):
    # Dirichlet outflow at x lo, Neumann at x hi; fine boxes touch the
    # periodic y seam from one side only (the seam face is held by one box)
    nc = (16, 8, 8)
    cboxes = B.chop_domain(nc, 8)
    fboxes = [((6, 0, 2), (13, 5, 9)), ((20, 10, 0), (25, 15, 15))]
    r = _run(orc, ctx, nc, cboxes, fboxes, (0, 1, 1), bclo=(BCD, 0, 0), bchi=(BCN, 0, 0), seed=2)
    _check(r)


def test_amr_mac_mixed_precision(orc, ctx):
    nc = (16, 12, 12)
    cboxes = B.chop_domain(nc, 8)
    fboxes = [((8, 4, 4), (15, 11, 11)), ((16, 4, 4), (23, 9, 7))]
    r = _run(orc, ctx, nc, cboxes, fboxes, (0, 1, 0), bclo=(BCD, 0, BCN), bchi=(BCN, 0, BCD),
             mode=1, seed=3)
    _check(r)


def test_amr_mac_errors(ctx):
    from incflo_amd import mgx
    nc = (8, 8, 8)
    cboxes = B.chop_domain(nc, 8)
    lev_c = mgx.make_level(nc, cboxes, periodic=(0, 1, 1))
    # more levels than the library holds
    with pytest.raises(mgx.MgxError):
        mgx.AmrMacProjector(ctx, [lev_c] * 5)
    # unaligned fine box
    lev_f = mgx.make_level((16, 16, 16), [((3, 2, 2), (8, 5, 5))], periodic=(0, 1, 1))
    with pytest.raises(mgx.MgxError):
        mgx.AmrMacProjector(ctx, [lev_c, lev_f])
    lev_f = mgx.make_level((16, 16, 16), [((4, 2, 2), (9, 5, 5))], periodic=(0, 1, 1))
    P = mgx.AmrMacProjector(ctx, [lev_c, lev_f])
    with pytest.raises(mgx.MgxError):
        P.setDomainBC((BCR, 0, 0), (BCN, 0, 0))  # Robin faces are not supported here
    # project before initProjector
    P.setDomainBC((BCD, 0, 0), (BCN, 0, 0))
    with pytest.raises(mgx.MgxError):
        P.project([[ctx.array(B.face_field_size(cboxes, 0, d)) for d in range(3)],
                   [ctx.array(B.face_field_size([((4, 2, 2), (9, 5, 5))], 0, d)) for d in range(3)]],
                  0, 1e-10)
    P.close()
