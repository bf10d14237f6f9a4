"""GPU parity of the composite nodal projection over three AMR levels
(mgx_amr_nodal_* with nlevels = 3) against the oracle (oracle.amr.AmrLevels),
through the C ABI.

The oracle solves the composite system with Jacobi-preconditioned conjugate
gradients, the GPU with composite V-cycles from the finest level down, so both
are driven to 1e-12 / 1e-13 and compared at 1e-9 relative: phi on every
level (on each fine level's closure), the projected velocity and G phi on
every level's cells (cells under a finer level averaged down from the finest
level).  Periodic, and x bounded with both fine levels touching the Dirichlet
and the Neumann face.  The mixed FP32/FP64 mode reaches the FP64 tolerance;
its phi is compared with the FP64 oracle at 1e-7."""
import sys

import numpy as np
import pytest

from incflo_amd import boxes as B
from oracle.amr import AmrLevels

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


PER = dict(per=(1, 1, 1), bclo=(0, 0, 0), bchi=(0, 0, 0),
           levels=[[(4, 4, 4, 23, 23, 19), (24, 8, 6, 27, 15, 13)],
                   [(10, 10, 10, 29, 25, 29), (30, 14, 12, 43, 29, 27), (44, 18, 14, 51, 27, 23)]])
BND = dict(per=(0, 1, 1), bclo=(1, 0, 0), bchi=(2, 0, 0),
           levels=[[(0, 4, 4, 15, 23, 19), (20, 6, 6, 31, 17, 15)],
                   [(0, 10, 10, 23, 33, 33), (44, 14, 14, 63, 29, 29)]])


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _run(ctx, case, mode, rtol, seed=3):
    import oracle
    from incflo_amd import mgx
    nc = (16, 16, 16)
    per = case["per"]
    gc = oracle.Geom.make(nc, periodic=per, bclo=case["bclo"], bchi=case["bchi"])
    A = AmrLevels(gc, case["levels"])
    rng = np.random.default_rng(seed)
    sigs = [0.01 / (1.0 + 0.5 * rng.random(g.cell_shape())) for g in A.g]
    us = [np.where(A.fmask[l][None], rng.standard_normal((3,) + g.cell_shape()), 0.0)
          for l, g in enumerate(A.g)]
    boxes = [B.chop_domain(nc, 8)] + [[(b[:3], b[3:]) for b in bl] for bl in case["levels"]]
    levs = [mgx.make_level(tuple(A.g[l].n), boxes[l], periodic=per) for l in range(3)]
    P = mgx.AmrNodalProjector(ctx, levs)
    if not all(per):
        P.setDomainBC(case["bclo"], case["bchi"])
    P.set_sigma([ctx.array(B.to_boxed(sigs[l], boxes[l], 0, periodic=per)) for l in range(3)], 0)
    vel = [ctx.array(B.to_boxed(us[l], boxes[l], 1, periodic=per)) for l in range(3)]
    phi = [ctx.array(B.field_size(boxes[l], 0, 1, True)) for l in range(3)]
    gp = [ctx.array(B.field_size(boxes[l], 0, 3)) for l in range(3)]
    info = P.project(vel, 1, rtol, 0.0, phis=phi, gphis=gp, mode=mode)
    g = dict(u=[B.from_boxed(vel[l].download(), boxes[l], 1, A.g[l].cell_shape(), ncomp=3)
                for l in range(3)],
             ph=[B.from_boxed(phi[l].download(), boxes[l], 0, A.g[l].node_shape(), nodal=True)
                 for l in range(3)],
             gp=[B.from_boxed(gp[l].download(), boxes[l], 0, A.g[l].cell_shape(), ncomp=3)
                 for l in range(3)])
    P.close()
    return A, sigs, us, info, g


@pytest.mark.parametrize("case", [PER, BND], ids=["periodic", "bounded"])
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_three_levels_match_oracle(ctx, orc, case, mode):
    from incflo_amd import mgx
    _say(f"case amr 3 levels {case['per']} {mode}")
    m = mgx.MODE_FP64 if mode == "fp64" else mgx.MODE_MIXED
    rtol = 1e-12 if mode == "fp64" else 1e-10
    A, sigs, us, info, g = _run(ctx, case, m, rtol)
    assert info.converged and info.rel_resid <= rtol
    _say(f"  composite V-cycles {info.iters}")
    assert info.iters <= (40 if mode == "fp64" else 30)
    un, ph, gp, _ = A.project(sigs, us, rtol=1e-13)
    tol = 1e-9 if mode == "fp64" else 1e-7
    assert _rel(g["ph"][0], ph[0]) <= tol
    for l in (1, 2):
        cl = A.ftype[l] != A.OUT
        assert _rel(g["ph"][l][cl], ph[l][cl]) <= tol, l
    for l in range(3):
        f = A.fmask[l]
        assert _rel(g["u"][l][:, f], un[l][:, f]) <= tol, l
        assert _rel(g["gp"][l][:, f], gp[l][:, f]) <= tol, l
    # cells under a finer level hold exactly the mean of their 8 children
    for l in (1, 0):
        c = A.cov[l]
        assert np.abs(g["u"][l][:, c] - A.average_down(g["u"][l + 1])[:, c]).max() < 1e-15


def test_levels_must_nest_and_stay_off_inflow_faces(ctx):
    from incflo_amd import mgx
    nc = (16, 16, 16)
    cb = B.chop_domain(nc, 8)
    lev0 = mgx.make_level(nc, cb)
    lev1 = mgx.make_level((32, 32, 32), [((4, 4, 4), (23, 23, 23))])
    bad2 = mgx.make_level((64, 64, 64), [((8, 8, 8), (31, 31, 31))])  # touches level 1's edge
    P = mgx.AmrNodalProjector(ctx, [lev0, lev1, bad2])
    with pytest.raises(mgx.MgxError) as e:
        P.set_sigma([ctx.array(B.field_size(cb, 0, 1))] * 3, 0)
    assert e.value.code == mgx.MGX_ERR_ARG and "nested" in str(e.value)
    P.close()
    levi = mgx.make_level(nc, cb, periodic=(0, 1, 1))
    fin = mgx.make_level((32, 32, 32), [((0, 4, 4), (7, 11, 11))], periodic=(0, 1, 1))
    P = mgx.AmrNodalProjector(ctx, [levi, fin])
    P.setDomainBC((3, 0, 0), (2, 0, 0))
    with pytest.raises(mgx.MgxError) as e:
        P.set_sigma([ctx.array(B.field_size(cb, 0, 1))] * 2, 0)
    assert e.value.code == mgx.MGX_ERR_UNSUPPORTED and "inflow" in str(e.value)
    P.close()


def test_one_handle_switches_precision(ctx, orc):
    """FP64, mixed and FP64 again on one handle (the V-cycle buffers change
    element type between the solves) give the oracle's projection each time."""
    import oracle
    from incflo_amd import mgx
    nc = (16, 16, 16)
    case = PER
    _say("case amr 3 levels, one handle, fp64 / mixed / fp64")
    gc = oracle.Geom.make(nc)
    A = AmrLevels(gc, case["levels"])
    rng = np.random.default_rng(5)
    sigs = [0.01 / (1.0 + 0.5 * rng.random(g.cell_shape())) for g in A.g]
    us = [np.where(A.fmask[l][None], rng.standard_normal((3,) + g.cell_shape()), 0.0)
          for l, g in enumerate(A.g)]
    boxes = [B.chop_domain(nc, 8)] + [[(b[:3], b[3:]) for b in bl] for bl in case["levels"]]
    levs = [mgx.make_level(tuple(A.g[l].n), boxes[l]) for l in range(3)]
    P = mgx.AmrNodalProjector(ctx, levs)
    P.set_sigma([ctx.array(B.to_boxed(sigs[l], boxes[l], 0)) for l in range(3)], 0)
    un, ph, _, _ = A.project(sigs, us, rtol=1e-13)
    for mode, rtol, tol in ((mgx.MODE_FP64, 1e-12, 1e-9), (mgx.MODE_MIXED, 1e-10, 1e-7),
                            (mgx.MODE_FP64, 1e-12, 1e-9)):
        vel = [ctx.array(B.to_boxed(us[l], boxes[l], 1)) for l in range(3)]
        info = P.project(vel, 1, rtol, 0.0, mode=mode)
        assert info.converged and info.rel_resid <= rtol
        for l in range(3):
            got = B.from_boxed(vel[l].download(), boxes[l], 1, A.g[l].cell_shape(), ncomp=3)
            f = A.fmask[l]
            assert _rel(got[:, f], un[l][:, f]) <= tol, (mode, l)
    P.close()
