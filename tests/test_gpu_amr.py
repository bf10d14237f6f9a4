"""GPU parity of the two-level composite nodal projection (mgx_amr_nodal_*)
against the oracle (oracle/amr.py), through the C ABI.

The oracle solves the composite system with Jacobi-preconditioned conjugate
gradients, the GPU with composite V-cycles, so both are driven to 1e-12 /
1e-13 and compared at 1e-9 relative (phi on both levels, the projected
velocity and G phi, with covered coarse cells averaged down).  The mixed
FP32/FP64 mode reaches the same FP64 tolerance: its phi is checked against the
FP64 one at 1e-7 relative and its composite residual with the oracle's
operator at 2e-10 of ||b||."""
import sys

import numpy as np
import pytest

from incflo_amd import boxes as B
from oracle.amr import AmrGeom

pytestmark = pytest.mark.gpu

D, N, INF = 1, 2, 3


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from incflo_amd import mgx
    c = mgx.Context(0)
    yield c
    c.close()


def _case(orc, nc, fboxes, periodic=(1, 1, 1), bclo=(0, 0, 0), bchi=(0, 0, 0), ph=(1.0, 1.0, 1.0),
          seed=1, inflow=None):
    import oracle
    gc = oracle.Geom.make(nc, periodic=periodic, bclo=bclo, bchi=bchi, prob_hi=ph)
    A = AmrGeom(gc, fboxes)
    rng = np.random.default_rng(seed)
    rho_c = 1.0 + 0.5 * rng.random(A.gc.cell_shape())
    rho_f = 1.0 + 0.5 * rng.random(A.gf.cell_shape())
    dt = 0.01
    uc = rng.standard_normal((3,) + A.gc.cell_shape())
    uf = np.where(A.fmask[None], rng.standard_normal((3,) + A.gf.cell_shape()), 0.0)
    return A, dt / rho_c, dt / rho_f, uc, uf


def _gpu_project(ctx, A, sig_c, sig_f, uc, uf, cboxes, fboxes, mode, rtol, periodic,
                 bclo=None, bchi=None, inflow=None):
    import oracle
    from incflo_amd import mgx
    gc, gf = A.gc, A.gf
    ph = tuple(gc.dx[d] * gc.n[d] for d in range(3))
    lev_c = mgx.make_level(tuple(gc.n), cboxes, periodic=periodic, prob_hi=ph)
    lev_f = mgx.make_level(tuple(gf.n), fboxes, periodic=periodic, prob_hi=ph)
    P = mgx.AmrNodalProjector(ctx, [lev_c, lev_f])
    if bclo is not None:
        P.setDomainBC(bclo, bchi)
    P.set_sigma([ctx.array(B.to_boxed(sig_c, cboxes, 0, periodic=periodic)),
                 ctx.array(B.to_boxed(sig_f, fboxes, 0, periodic=periodic))], 0)
    # coarse velocity with its domain ghost cells (inflow values) as the caller fills them
    ug = np.zeros(gc.ghost_vec_shape())
    ug[:, 1:-1, 1:-1, 1:-1] = uc
    oracle.fill_vel_ghosts(gc, ug, inflow)
    vc_h = B.to_boxed(uc, cboxes, 1, periodic=periodic)
    if not all(periodic):
        vc_h = B.set_ghosts_from(vc_h, cboxes, 1, ug)
    vc = ctx.array(vc_h)
    vf = ctx.array(B.to_boxed(uf, fboxes, 1, periodic=periodic))
    pc = ctx.array(B.field_size(cboxes, 0, 1, True))
    pf = ctx.array(B.field_size(fboxes, 0, 1, True))
    gpc = ctx.array(B.field_size(cboxes, 0, 3))
    gpf = ctx.array(B.field_size(fboxes, 0, 3))
    info = P.project([vc, vf], 1, rtol, 0.0, phis=[pc, pf], gphis=[gpc, gpf], mode=mode)
    out = dict(
        uc=B.from_boxed(vc.download(), cboxes, 1, gc.cell_shape(), ncomp=3),
        uf=B.from_boxed(vf.download(), fboxes, 1, gf.cell_shape(), ncomp=3),
        phc=B.from_boxed(pc.download(), cboxes, 0, gc.node_shape(), nodal=True),
        phf=B.from_boxed(pf.download(), fboxes, 0, gf.node_shape(), nodal=True),
        gpc=B.from_boxed(gpc.download(), cboxes, 0, gc.cell_shape(), ncomp=3),
        gpf=B.from_boxed(gpf.download(), fboxes, 0, gf.cell_shape(), ncomp=3))
    P.close()
    return info, out


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


FINE = [(8, 8, 8, 23, 15, 15), (8, 16, 8, 15, 23, 15), (24, 2, 20, 31, 9, 27)]


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_amr_periodic_matches_oracle(ctx, orc, mode):
    from incflo_amd import mgx
    _say(f"case amr periodic {mode}")
    nc = (16, 16, 16)
    fboxes = [(b[:3], b[3:]) for b in FINE]
    A, sig_c, sig_f, uc, uf = _case(orc, nc, FINE)
    cboxes = B.chop_domain(nc, 8)
    m = mgx.MODE_FP64 if mode == "fp64" else mgx.MODE_MIXED
    rtol = 1e-12 if mode == "fp64" else 1e-10
    info, g = _gpu_project(ctx, A, sig_c, sig_f, uc, uf, cboxes, fboxes, m, rtol, (1, 1, 1))
    assert info.converged and info.rel_resid <= rtol
    _say(f"  composite V-cycles {info.iters}")
    ucn, ufn, phc, phf, gpc, gpf, its = A.project(sig_c, sig_f, uc, uf, rtol=1e-13)
    closure = A.ftype != A.OUT
    tol = 1e-9 if mode == "fp64" else 1e-7
    assert _rel(g["phc"], phc) <= tol
    assert _rel(g["phf"][closure], phf[closure]) <= tol
    assert _rel(g["uc"], ucn) <= tol
    assert _rel(g["uf"][:, A.fmask], ufn[:, A.fmask]) <= tol
    assert _rel(g["gpc"], gpc) <= tol
    assert _rel(g["gpf"][:, A.fmask], gpf[:, A.fmask]) <= tol
    if mode == "mixed":
        # composite residual of the mixed answer with the oracle's FP64 operator
        bc, bf = A.divergence(uc, uf)
        dof_f = A.ftype == A.DOF
        mu = (8 * bc.sum() + bf.sum()) / (A.cdof.sum() + dof_f.sum())
        yc, yf = A.apply(sig_c, sig_f, np.where(A.cdof, g["phc"], 0.0),
                         np.where(dof_f, g["phf"], 0.0))
        r = max(np.abs(np.where(A.cdof, bc - mu / 8 - yc, 0.0)).max(),
                np.abs(np.where(dof_f, bf - mu - yf, 0.0)).max())
        assert r <= 2e-10 * max(np.abs(bc).max(), np.abs(bf).max())


def test_amr_channel_matches_oracle(ctx, orc):
    """inflow at x-lo, outflow (Dirichlet) at x-hi, walls in y, periodic in z;
    the fine region stays off the non-periodic faces"""
    from incflo_amd import mgx
    _say("case amr channel")
    nc = (16, 8, 8)
    per, bclo, bchi = (0, 0, 1), (INF, N, 0), (D, N, 0)
    fine = [(6, 4, 0, 17, 11, 7), (18, 4, 4, 21, 7, 15)]
    inflow = np.zeros((6, 3))
    inflow[0] = [1.0, 0.0, 0.0]
    A, sig_c, sig_f, uc, uf = _case(orc, nc, fine, periodic=per, bclo=bclo, bchi=bchi,
                                    ph=(2.0, 1.0, 1.0), seed=3)
    cboxes = B.chop_domain(nc, 8)
    fboxes = [(b[:3], b[3:]) for b in fine]
    info, g = _gpu_project(ctx, A, sig_c, sig_f, uc, uf, cboxes, fboxes, mgx.MODE_FP64, 1e-12, per,
                           bclo, bchi, inflow)
    assert info.converged
    ucn, ufn, phc, phf, gpc, gpf, its = A.project(sig_c, sig_f, uc, uf, rtol=1e-13,
                                                  inflow_vel=inflow)
    closure = A.ftype != A.OUT
    assert _rel(g["phc"], phc) <= 1e-9
    assert _rel(g["phf"][closure], phf[closure]) <= 1e-9
    assert _rel(g["uc"], ucn) <= 1e-9
    assert _rel(g["uf"][:, A.fmask], ufn[:, A.fmask]) <= 1e-9


def test_amr_rejects_bad_levels(ctx):
    from incflo_amd import mgx
    _say("case amr argument checks")
    nc = (16, 16, 16)
    lev_c = mgx.make_level(nc, B.chop_domain(nc, 8))
    bad = mgx.make_level((32, 32, 32), [((1, 0, 0), (8, 7, 7))])
    with pytest.raises(mgx.MgxError) as e:
        mgx.AmrNodalProjector(ctx, [lev_c, bad])
    assert e.value.code == mgx.MGX_ERR_ARG
    # a fine box on a non-periodic face: accepted for walls and outflow, refused
    # once that face is an inflow face
    wall = mgx.make_level((32, 16, 16), [((0, 0, 0), (7, 7, 7))], periodic=(0, 1, 1))
    lev_w = mgx.make_level((16, 8, 8), B.chop_domain((16, 8, 8), 8), periodic=(0, 1, 1))
    P = mgx.AmrNodalProjector(ctx, [lev_w, wall])
    sig = [ctx.array(B.field_size(B.chop_domain((16, 8, 8), 8), 0, 1)),
           ctx.array(B.field_size([((0, 0, 0), (7, 7, 7))], 0, 1))]
    P.setDomainBC((D, 0, 0), (N, 0, 0))
    P.set_sigma(sig, 0)
    P.setDomainBC((INF, 0, 0), (N, 0, 0))
    with pytest.raises(mgx.MgxError) as e:
        P.set_sigma(sig, 0)
    assert e.value.code == mgx.MGX_ERR_UNSUPPORTED
    P.close()
    # more levels than the library holds
    with pytest.raises(mgx.MgxError) as e:
        mgx.AmrNodalProjector(ctx, [lev_c] * 5)
    assert e.value.code == mgx.MGX_ERR_UNSUPPORTED


def test_average_down_faces_and_cells_match_oracle(ctx):
    """average_down_faces / average_down through the C ABI against the oracle,
    on a fine level of two boxes (one on the periodic seam side)."""
    from incflo_amd import mgx
    from oracle import amr as A
    print("case average down", flush=True)
    nc, nf = (16, 16, 8), (32, 32, 16)
    cb = B.chop_domain(nc, 8)
    fb = [((4, 8, 0), (11, 19, 7)), ((20, 2, 4), (31, 9, 15))]
    lc = mgx.make_level(nc, cb)
    lf = mgx.make_level(nf, fb)
    F = np.zeros((nf[2], nf[1], nf[0]), bool)
    for lo, hi in fb:
        F[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
    rng = np.random.default_rng(41)
    fshape = lambda n, d: tuple(n[2 - a] + (1 if a == 2 - d else 0) for a in range(3))
    ff = [rng.standard_normal((2,) + fshape(nf, d)) for d in range(3)]
    fc = [rng.standard_normal((2,) + fshape(nc, d)) for d in range(3)]
    for d in range(3):  # periodic: the duplicate coarse face repeats face 0
        ax = 3 - d
        sl0 = [slice(None)] * 4; sl0[ax] = 0
        sl1 = [slice(None)] * 4; sl1[ax] = -1
        fc[d][tuple(sl1)] = fc[d][tuple(sl0)]
    fd = [ctx.array(B.faces_to_boxed_n(ff[d], fb, 0, d)) for d in range(3)]
    cd = [ctx.array(B.faces_to_boxed_n(fc[d], cb, 0, d)) for d in range(3)]
    mgx.amr_average_down_faces(ctx, lc, lf, fd, 0, cd, 0, ncomp=2)
    for d in range(3):
        got = B.faces_from_boxed_n(cd[d].download(), cb, 0, d, fshape(nc, d), 2)
        for q in range(2):
            ref = A.average_down_faces(F, [ff[e][q] for e in range(3)], [fc[e][q] for e in range(3)])[d]
            assert np.max(np.abs(got[q] - ref)) <= 1e-15 * np.max(np.abs(ref)), (d, q)
    uf = rng.standard_normal((3, nf[2], nf[1], nf[0]))
    uc = rng.standard_normal((3, nc[2], nc[1], nc[0]))
    ufd = ctx.array(B.to_boxed(uf, fb, 1))
    ucd = ctx.array(B.to_boxed(uc, cb, 1))
    mgx.amr_average_down(ctx, lc, lf, ufd, 1, ucd, 1, ncomp=3)
    got = B.from_boxed(ucd.download(), cb, 1, (nc[2], nc[1], nc[0]), ncomp=3)
    for q in range(3):
        ref = A.average_down_cells(F, uf[q], uc[q])
        assert np.max(np.abs(got[q] - ref)) <= 1e-15 * np.max(np.abs(ref))


def test_amr_coefficients_survive_resetup(ctx, orc):
    """set_domain_bc / set_options after the coefficients (incflo's order) set
    the levels up again and keep sigma / beta: the result equals the one
    without the extra call."""
    from incflo_amd import mgx
    _say("case amr coefficients kept across set-up")
    nc = (16, 16, 16)
    fboxes = [(b[:3], b[3:]) for b in FINE]
    A, sig_c, sig_f, uc, uf = _case(orc, nc, FINE, seed=4)
    cboxes = B.chop_domain(nc, 8)
    lev_c = mgx.make_level(nc, cboxes)
    lev_f = mgx.make_level(tuple(2 * v for v in nc), fboxes)
    outs = []
    for extra in (False, True):
        P = mgx.AmrNodalProjector(ctx, [lev_c, lev_f])
        P.set_sigma([ctx.array(B.to_boxed(sig_c, cboxes, 0)), ctx.array(B.to_boxed(sig_f, fboxes, 0))], 0)
        if extra:
            P.set_options(mgx.default_opts())
            P.setDomainBC((0, 0, 0), (0, 0, 0))
        v = [ctx.array(B.to_boxed(uc, cboxes, 0)), ctx.array(B.to_boxed(uf, fboxes, 0))]
        info = P.project(v, 0, 1e-12, 0.0)
        assert info.converged
        outs.append([a.download() for a in v])
        P.close()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
    # MAC: beta kept across set_options
    rng = np.random.default_rng(9)
    gc = orc.Geom.make(nc)
    gf = orc.Geom.make(tuple(2 * v for v in nc))
    bc = [0.5 + rng.random(orc.face_shape(gc, d)) for d in range(3)]
    bf = [0.5 + rng.random(orc.face_shape(gf, d)) for d in range(3)]
    for d in range(3):
        ax = 2 - d
        for b in (bc[d], bf[d]):
            sl0 = [slice(None)] * 3
            sln = [slice(None)] * 3
            sl0[ax] = 0
            sln[ax] = -1
            b[tuple(sln)] = b[tuple(sl0)]
    u0 = [[rng.standard_normal(orc.face_shape(gc, d)) for d in range(3)],
          [rng.standard_normal(orc.face_shape(gf, d)) for d in range(3)]]
    outs = []
    for extra in (False, True):
        M = mgx.AmrMacProjector(ctx, [lev_c, lev_f])
        M.initProjector([[ctx.array(B.faces_to_boxed(bc[d], cboxes, 0, d)) for d in range(3)],
                         [ctx.array(B.faces_to_boxed(bf[d], fboxes, 0, d)) for d in range(3)]], 0)
        if extra:
            M.set_options(mgx.default_opts(maxiter=200, bottom_maxiter=200))
        ud = [[ctx.array(B.faces_to_boxed(u0[0][d], cboxes, 0, d)) for d in range(3)],
              [ctx.array(B.faces_to_boxed(u0[1][d], fboxes, 0, d)) for d in range(3)]]
        info = M.project(ud, 0, 1e-12)
        assert info.converged
        outs.append([a.download() for lev in ud for a in lev])
        M.close()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
