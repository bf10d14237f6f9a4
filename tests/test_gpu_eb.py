"""GPU parity of the embedded-boundary nodal projection against the oracle
(oracle/orc_eb.c), through the C ABI, on a channel with a cylinder (the
configuration of the EB benchmark case, at test size): inflow at x-lo,
outflow (Dirichlet) at x-hi, walls in y, periodic in z.

The oracle solves with Jacobi-preconditioned conjugate gradients, the GPU
with cut-cell multigrid, so both are driven to 1e-12 and compared at 1e-9
relative (phi, velocity); the operator apply agrees at rounding level
(1e-12 relative)."""
import sys

import numpy as np
import pytest

from incflo_amd import boxes as B

pytestmark = pytest.mark.gpu

D, N, INF = 1, 2, 3
CX, CY, R = 0.5, 0.5, 0.2


@pytest.fixture(scope="module")
def ctx():
    from incflo_amd import mgx
    c = mgx.Context(0)
    yield c
    c.close()


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _setup(ctx, orc, n=(32, 16, 16), ph=(2.0, 1.0, 1.0), mgs=(16, 8, 8)):
    from incflo_amd import mgx
    per, bclo, bchi = (0, 0, 1), (INF, N, 0), (D, N, 0)
    g = orc.Geom.make(n, periodic=per, bclo=bclo, bchi=bchi, prob_hi=ph)
    bxs = B.chop_domain(n, mgs)
    P = mgx.NodalProjector(ctx, mgx.make_level(n, bxs, periodic=per, prob_hi=ph))
    P.setDomainBC(bclo, bchi)
    P.set_eb_cylinder(CX, CY, R)
    ct, mom = orc.eb_cylinder(g, CX, CY, R)
    return g, bxs, P, ct, mom


def test_eb_apply_matches_oracle(ctx, orc):
    _say("case eb apply")
    g, bxs, P, ct, mom = _setup(ctx, orc)
    assert (ct == 2).sum() > 0 and (ct == 0).sum() > 0
    rng = np.random.default_rng(1)
    sigma = 0.5 + rng.random(g.cell_shape())
    phi = rng.standard_normal(g.node_shape())
    mask = orc.eb_node_mask(g, ct).astype(bool)
    P.set_sigma(ctx.array(B.to_boxed(sigma, bxs, 0)), 0)
    pd = ctx.array(B.to_boxed(phi, bxs, 0, nodal=True, periodic=(0, 0, 1)))
    od = ctx.array(B.field_size(bxs, 0, 1, True))
    P.apply(pd, 0, od, 0)
    got = B.from_boxed(od.download(), bxs, 0, g.node_shape(), nodal=True)
    ref = orc.eb_nodal_apply(g, ct, mom, sigma, np.where(mask, phi, 0.0))
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert np.all(got[~mask] == 0.0)
    P.close()


def test_eb_solve_matches_oracle(ctx, orc):
    from incflo_amd import mgx
    _say("case eb solve")
    g, bxs, P, ct, mom = _setup(ctx, orc)
    rng = np.random.default_rng(2)
    sigma = 0.5 + rng.random(g.cell_shape())
    mask = orc.eb_node_mask(g, ct).astype(bool)
    b = np.where(mask, rng.standard_normal(g.node_shape()), 0.0)
    P.set_sigma(ctx.array(B.to_boxed(sigma, bxs, 0)), 0)
    rd = ctx.array(B.to_boxed(b, bxs, 0, nodal=True, periodic=(0, 0, 1)))
    xd = ctx.array(B.field_size(bxs, 0, 1, True))
    info = P.solve(xd, 0, rd, 0, 1e-12, 0.0, mode=mgx.MODE_FP64)
    assert info.converged and info.rel_resid <= 1e-12
    x = B.from_boxed(xd.download(), bxs, 0, g.node_shape(), nodal=True)
    xo, its, ret = orc.eb_nodal_solve(g, ct, mom, sigma, b, rtol=1e-13)
    assert ret == 0
    assert np.max(np.abs(x - xo)) <= 1e-9 * np.max(np.abs(xo))
    # the residual of the GPU answer measured by the oracle's operator
    r = np.where(mask, b - orc.eb_nodal_apply(g, ct, mom, sigma, x), 0.0)
    assert np.max(np.abs(r)) <= 2e-12 * np.max(np.abs(b))
    with pytest.raises(mgx.MgxError) as e:
        P.solve(xd, 0, rd, 0, 1e-10, 0.0, mode=mgx.MODE_MIXED)
    assert e.value.code == mgx.MGX_ERR_UNSUPPORTED
    P.close()


def test_eb_nodal_projection_channel(ctx, orc):
    from incflo_amd import mgx
    _say("case eb projection")
    g, bxs, P, ct, mom = _setup(ctx, orc, n=(64, 32, 16), ph=(2.0, 1.0, 0.5), mgs=(16, 16, 8))
    rng = np.random.default_rng(3)
    u = np.zeros((3,) + g.cell_shape())
    u[0] = 1.0
    u += 0.1 * rng.standard_normal(u.shape)
    u *= (ct > 0)
    rho = 1.0 + 0.2 * rng.random(g.cell_shape())
    s = 0.01
    inflow = np.zeros((6, 3))
    inflow[0] = [1.0, 0.0, 0.0]
    vd = ctx.array(B.to_boxed(u, bxs, 1, periodic=(0, 0, 1)))
    gpd = ctx.array(B.field_size(bxs, 0, 3))
    pd = ctx.array(B.field_size(bxs, 0, 1, True))
    rd = ctx.array(B.to_boxed(rho, bxs, 0))
    info = P.apply_nodal_projection(vd, 1, gpd, 0, pd, 0, s, density=rd, rho_ng=0,
                                    inflow_vel=inflow, rtol=1e-12, atol=0.0,
                                    mode=mgx.MODE_FP64)
    assert info.converged
    unew = B.from_boxed(vd.download(), bxs, 1, g.cell_shape(), ncomp=3)
    gp = B.from_boxed(gpd.download(), bxs, 0, g.cell_shape(), ncomp=3)
    p = B.from_boxed(pd.download(), bxs, 0, g.node_shape(), nodal=True)
    # oracle composition: ghosts, EB divergence, CG solve, EB gradient, update
    ug = np.zeros(g.ghost_vec_shape())
    ug[:, 1:-1, 1:-1, 1:-1] = u
    orc.fill_vel_ghosts(g, ug, inflow)
    b = orc.eb_nodal_divergence(g, ct, mom, ug)
    sigma = s / rho
    phi, its, ret = orc.eb_nodal_solve(g, ct, mom, sigma, b, rtol=1e-13)
    assert ret == 0
    G = orc.eb_nodal_gradient(g, ct, mom, phi)
    uref = (u - sigma[None] * G) * (ct > 0)
    assert np.max(np.abs(p - phi)) <= 1e-9 * np.max(np.abs(phi))
    assert np.max(np.abs(gp - G)) <= 1e-9 * np.max(np.abs(G))
    assert np.max(np.abs(unew - uref)) <= 1e-9 * np.max(np.abs(uref))
    assert np.all(unew[:, ct == 0] == 0.0)
    _say(f"  V-cycles {info.iters}, CG iterations {its}")
    P.close()
