"""GPU parity of the cell-centred viscous operator and Helmholtz solve
(mgx_diff_*) against the diffusion oracle (oracle/diffusion.py), through the
C ABI: DiffusionScalarOp's component solves (src/diffusion/DiffusionScalarOp.cpp:332-480)
and compute_divtau / compute_laps (:482-760).

Bars: FP64 solves and the explicit operator match the oracle to 1e-10
relative; the mixed (FP32 V-cycle in FP64 refinement) solve reaches the same
tolerance and the same answer to 1e-9."""
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


def _ghosted(rng, ncomp, n):
    nx, ny, nz = n
    return rng.standard_normal((ncomp, nz + 2, ny + 2, nx + 2))


def _planes(g, ug, bcs):
    """boundary values from the ghost layer of a ghosted canonical array, for
    the Dirichlet sides listed in bcs[(d, side)]"""
    out = {}
    for (d, side) in bcs:
        ax = 2 - d
        idx = 0 if side == 0 else ug.shape[ax] - 1
        pl = np.take(ug, idx, axis=ax)
        out[(d, side)] = pl[1:-1, 1:-1]
    return out


@pytest.mark.parametrize("mode", [0, 1])
def test_periodic_helmholtz_variable_coefficients(ctx, orc, mode):
    from incflo_amd import mgx
    from oracle import diffusion as D
    n = (24, 20, 16)
    _say(f"case periodic Helmholtz n={n} mode={mode}")
    g = orc.Geom.make(n, prob_hi=(1.2, 1.0, 0.8))
    bxs = B.chop_domain(n, 12)  # uneven boxes: 12 and 8 cells
    lev = mgx.make_level(n, bxs, prob_hi=(1.2, 1.0, 0.8))
    rng = np.random.default_rng(11)
    shape = g.cell_shape()
    rho = 1.0 + 0.5 * rng.random(shape)
    eta = 0.5 + rng.random(shape)
    u = rng.standard_normal((3,) + shape)
    dt = 0.02
    Dh = mgx.Diffusion(ctx, lev)
    Dh.set_eta(ctx.array(B.to_boxed(eta, bxs, 1)), 1)
    ud = ctx.array(B.to_boxed(u, bxs, 2))
    rd = ctx.array(B.to_boxed(rho, bxs, 0))
    info = Dh.solve(ud, 2, 3, dt, acoef=rd, rtol=1e-12, atol=0.0, mode=mode)
    assert info.converged
    got = B.from_boxed(ud.download(), bxs, 2, shape, ncomp=3)
    ref = D.diffuse_velocity(g, u, rho, eta, dt)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    _say(f"  vcycles {info.vcycles} err {err:.2e}")
    assert err <= (1e-10 if mode == 0 else 1e-9)
    # explicit operator (divtau = div(eta grad u) / rho)
    out = ctx.array(B.field_size(bxs, 0, 3))
    ud2 = ctx.array(B.to_boxed(u, bxs, 1))
    Dh.apply(ud2, 1, 3, out, 0, rho=rd)
    dtau = B.from_boxed(out.download(), bxs, 0, shape, ncomp=3)
    ref_t = D.divtau(g, u, rho, eta)
    assert np.max(np.abs(dtau - ref_t)) <= 1e-10 * np.max(np.abs(ref_t))
    Dh.close()


def test_dirichlet_neumann_sides_per_component(ctx, orc):
    """walls in x and z with values in the ghost cells; component 0 Dirichlet
    on both, components 1 and 2 Dirichlet in z and Neumann in x (the slip-wall
    pattern of get_diffuse_velocity_bc, src/diffusion/incflo_diffusion.cpp:187-195)"""
    from incflo_amd import mgx
    from oracle import BC_DIRICHLET, BC_NEUMANN
    from oracle import diffusion as D
    n = (16, 12, 20)
    _say(f"case bounded Helmholtz n={n}")
    per = (0, 1, 0)
    bxs = B.chop_domain(n, 8)
    lev = mgx.make_level(n, bxs, periodic=per)
    rng = np.random.default_rng(12)
    ug = _ghosted(rng, 3, n)
    shape = (n[2], n[1], n[0])
    rho = 1.0 + rng.random(shape)
    eta = 0.2 + rng.random(shape)
    dt = 0.05
    boxed = B.set_ghosts_from(np.zeros(B.field_size(bxs, 1, 3)), bxs, 1, ug, ncomp=3, gw=1)
    Dh = mgx.Diffusion(ctx, lev)
    Dh.setDomainBC((1, 0, 1), (1, 0, 1), comp=0)
    for c in (1, 2):
        Dh.setDomainBC((2, 0, 1), (2, 0, 1), comp=c)
    Dh.set_eta(ctx.array(B.to_boxed(eta, bxs, 0)), 0)
    ud = ctx.array(boxed)
    rd = ctx.array(B.to_boxed(rho, bxs, 0))
    info = Dh.solve(ud, 1, 3, dt, acoef=rd, rtol=1e-12, atol=0.0)
    assert info.converged
    got = B.from_boxed(ud.download(), bxs, 1, shape, ncomp=3)
    u = ug[:, 1:-1, 1:-1, 1:-1]
    bf = D.eta_faces(eta, per)
    for c in range(3):
        if c == 0:
            g = orc.Geom.make(n, periodic=per, bclo=(BC_DIRICHLET, 0, BC_DIRICHLET),
                              bchi=(BC_DIRICHLET, 0, BC_DIRICHLET))
            sides = [(0, 0), (0, 1), (2, 0), (2, 1)]
        else:
            g = orc.Geom.make(n, periodic=per, bclo=(BC_NEUMANN, 0, BC_DIRICHLET),
                              bchi=(BC_NEUMANN, 0, BC_DIRICHLET))
            sides = [(2, 0), (2, 1)]
        bv = _planes(g, ug[c], sides)
        ref = D.helmholtz_solve(g, rho, bf, dt, rho * u[c], bv)
        err = np.max(np.abs(got[c] - ref)) / np.max(np.abs(ref))
        _say(f"  comp {c} err {err:.2e}")
        assert err <= 1e-10
        # the explicit operator with the same boundary values
        out = ctx.array(B.field_size(bxs, 0, 3))
        Dh.apply(ctx.array(boxed), 1, 3, out, 0)
        L = B.from_boxed(out.download(), bxs, 0, shape, ncomp=3)[c]
        refL = D.apply(g, bf, u[c], bv)
        assert np.max(np.abs(L - refL)) <= 1e-10 * np.max(np.abs(refL))
    Dh.close()


def test_diffusion_errors(ctx):
    from incflo_amd import mgx
    n = (8, 8, 8)
    bxs = B.chop_domain(n, 8)
    Dh = mgx.Diffusion(ctx, mgx.make_level(n, bxs, periodic=(0, 1, 1)))
    u = ctx.array(B.field_size(bxs, 1, 3))
    with pytest.raises(mgx.MgxError) as e:  # no eta yet
        Dh.solve(u, 1, 3, 0.1)
    assert e.value.code == mgx.MGX_ERR_STATE
    with pytest.raises(mgx.MgxError) as e:  # periodic flag must match the geometry
        Dh.setDomainBC((0, 0, 0), (1, 0, 0))
    assert e.value.code == mgx.MGX_ERR_ARG
    with pytest.raises(mgx.MgxError) as e:
        Dh.setDomainBC((mgx.BC_ROBIN, 0, 0), (1, 0, 0))
    assert e.value.code == mgx.MGX_ERR_UNSUPPORTED
    Dh.close()
