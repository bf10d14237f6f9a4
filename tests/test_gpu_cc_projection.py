"""GPU parity of the cell-centred projection (mgx_apply_cc_projection,
incflo::ApplyCCProjection, src/projection/incflo_apply_cc_projection.cpp:199-487)
against the oracle (oracle/cc_projection.py) through the C ABI.  Bar: FP64
u, gp and p_cc to 1e-10 relative."""
import sys

import numpy as np
import pytest

from incflo_amd import boxes as B

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


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("incremental", [False, True])
def test_periodic_variable_density(ctx, orc, incremental):
    from incflo_amd import mgx
    from oracle import cc_projection as CC
    n = (24, 20, 16)
    _say(f"case cc projection periodic n={n} incremental={incremental}")
    g = orc.Geom.make(n)
    bxs = B.chop_domain(n, 12)
    lev = mgx.make_level(n, bxs)
    rng = np.random.default_rng(21 + incremental)
    shape = g.cell_shape()
    u = rng.standard_normal((3,) + shape)
    uo = rng.standard_normal((3,) + shape)
    gp = rng.standard_normal((3,) + shape)
    pc = rng.standard_normal(shape)
    rho = 1.0 + 0.5 * rng.random(shape)
    s = 0.02
    M = mgx.MacProjector(ctx, lev)
    vd = ctx.array(B.to_boxed(u, bxs, 1))
    vod = ctx.array(B.to_boxed(uo, bxs, 1))
    gpd = ctx.array(B.to_boxed(gp, bxs, 0))
    pd = ctx.array(B.to_boxed(pc, bxs, 0))
    rd = ctx.array(B.to_boxed(rho, bxs, 1))
    info = M.apply_cc_projection(vd, 1, gpd, 0, pd, 0, s, density=rd, rho_ng=1, vel_old=vod,
                                 vel_old_ng=1, incremental=incremental, rtol=1e-12, atol=0.0)
    assert info.converged
    un, dgp, phi, _ = CC.cc_project(g, u, gp, rho, s, vel_old=uo, incremental=incremental,
                                    rtol=1e-12)
    got_u = B.from_boxed(vd.download(), bxs, 1, shape, ncomp=3)
    got_gp = B.from_boxed(gpd.download(), bxs, 0, shape, ncomp=3)
    got_p = B.from_boxed(pd.download(), bxs, 0, shape)
    ref_gp = gp + dgp if incremental else dgp
    ref_p = pc + phi if incremental else phi
    eu, eg = _rel(got_u, un), _rel(got_gp, ref_gp)
    # phi of a periodic (singular) solve is defined up to a constant
    ep = _rel(got_p - got_p.mean(), ref_p - ref_p.mean())
    _say(f"  iters {info.iters} err u {eu:.2e} gp {eg:.2e} p {ep:.2e}")
    assert eu <= 1e-10 and eg <= 1e-10 and ep <= 1e-10
    M.close()


def test_channel_with_inflow(ctx, orc):
    from incflo_amd import mgx
    from oracle import BC_DIRICHLET, BC_INFLOW, BC_NEUMANN
    from oracle import cc_projection as CC
    n = (32, 16, 16)
    _say(f"case cc projection channel n={n}")
    per = (0, 0, 1)
    lo, hi = (BC_INFLOW, BC_NEUMANN, 0), (BC_DIRICHLET, BC_NEUMANN, 0)
    g = orc.Geom.make(n, periodic=per, bclo=lo, bchi=hi, prob_hi=(2.0, 1.0, 1.0))
    bxs = B.chop_domain(n, 16)
    lev = mgx.make_level(n, bxs, periodic=per, prob_hi=(2.0, 1.0, 1.0))
    rng = np.random.default_rng(31)
    shape = g.cell_shape()
    u = rng.standard_normal((3,) + shape)
    u[0] += 1.0
    gp = 0.1 * rng.standard_normal((3,) + shape)
    pc = np.zeros(shape)
    inflow = np.zeros((6, 3))
    inflow[0] = (1.0, 0.1, -0.05)
    M = mgx.MacProjector(ctx, lev)
    M.setDomainBC(lo, hi)
    vd = ctx.array(B.to_boxed(u, bxs, 1, periodic=per))
    gpd = ctx.array(B.to_boxed(gp, bxs, 0))
    pd = ctx.array(B.to_boxed(pc, bxs, 0))
    um = [ctx.array(B.face_field_size(bxs, 0, d)) for d in range(3)]
    info = M.apply_cc_projection(vd, 1, gpd, 0, pd, 0, 0.05, const_rho=1.0, inflow_vel=inflow,
                                 umac=um, rtol=1e-12, atol=0.0)
    assert info.converged
    un, gpn, phi, _ = CC.cc_project(g, u, gp, None, 0.05, inflow=inflow, const_rho=1.0,
                                    rtol=1e-12)
    got_u = B.from_boxed(vd.download(), bxs, 1, shape, ncomp=3)
    got_gp = B.from_boxed(gpd.download(), bxs, 0, shape, ncomp=3)
    got_p = B.from_boxed(pd.download(), bxs, 0, shape)
    eu, eg, ep = _rel(got_u, un), _rel(got_gp, gpn), _rel(got_p, phi)
    _say(f"  iters {info.iters} err u {eu:.2e} gp {eg:.2e} p {ep:.2e}")
    assert eu <= 1e-10 and eg <= 1e-10 and ep <= 1e-10
    # the projected MAC velocities are divergence free to the solver tolerance
    fs = [B.faces_from_boxed(um[d].download(), bxs, 0, d, orc.face_shape(g, d)) for d in range(3)]
    div = orc.mac_divergence(g, fs)
    assert np.max(np.abs(div)) <= 1e-9 * np.max(np.abs(u)) * n[0]
    M.close()
