"""GPU parity of the MAC projection against the FP64 CPU oracle (through the C ABI).

Tolerances: FP64 solve / projection 1e-10 relative in phi and u_mac (both
sides run the same red-black V-cycle); mixed solve: relative residual <= 1e-10
by the oracle's operator; after the projection the discrete divergence is
the solver residual (<= rtol * ||div u||).
"""
import sys

import numpy as np
import pytest

from incflo_amd import boxes as B

pytestmark = pytest.mark.gpu

D_, N_, INF_ = 1, 2, 3


@pytest.fixture(scope="module")
def ctx():
    from incflo_amd import mgx
    c = mgx.Context(0)
    yield c
    c.close()


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _periodic_consistent(orc, g, arrs):
    for d in range(3):
        if g.periodic[d]:
            ax = 2 - d
            i0 = [slice(None)] * 3; i0[ax] = 0
            i1 = [slice(None)] * 3; i1[ax] = -1
            arrs[d][tuple(i1)] = arrs[d][tuple(i0)]
    return arrs


CASES = {
    "periodic": ((32, 32, 32), (1, 1, 1), (0, 0, 0), (0, 0, 0), (1, 1, 1)),
    "walls_dirichlet": ((32, 16, 24), (0, 0, 0), (D_, N_, N_), (N_, N_, D_), (2.0, 1.0, 1.5)),
    "channel": ((48, 16, 16), (0, 1, 1), (INF_, 0, 0), (D_, 0, 0), (3.0, 1.0, 1.0)),
}


def _mac(ctx, n, bxs, per, bclo, bchi, ph):
    from incflo_amd import mgx
    P = mgx.MacProjector(ctx, mgx.make_level(n, bxs, periodic=per, prob_hi=ph))
    if not all(per):
        P.setDomainBC(bclo, bchi)
    return P


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("mode", [0, 1])
def test_mac_project_matches_oracle(ctx, orc, case, mode):
    from incflo_amd import mgx
    n, per, bclo, bchi, ph = CASES[case]
    _say(f"case mac project {case} mode={mode}")
    g = orc.Geom.make(n, periodic=per, bclo=bclo, bchi=bchi, prob_hi=ph)
    rng = np.random.default_rng(7)
    rho = 1.0 + 0.5 * np.tanh(rng.standard_normal(g.cell_shape()))
    u = _periodic_consistent(orc, g, [rng.standard_normal(orc.face_shape(g, d)) for d in range(3)])
    if bclo[0] == INF_ and not per[0]:
        u[0][:, :, 0] = 1.0
    dt = 0.01
    bxs = B.chop_domain(n, (16, 8, 12))
    P = _mac(ctx, n, bxs, per, bclo, bchi, ph)
    P.set_beta_from_density(ctx.array(B.to_boxed(rho, bxs, 1, periodic=per)), 1, dt)
    ud = [ctx.array(B.faces_to_boxed(u[d], bxs, 0, d)) for d in range(3)]
    phid = ctx.array(B.field_size(bxs, 1, 1))
    rtol = 1e-12 if mode == mgx.MODE_FP64 else 1e-10
    info = P.project(ud[0], ud[1], ud[2], 0, rtol, 0.0, phi=phid, phi_ng=1, mode=mode)
    assert info.converged
    un = [B.faces_from_boxed(ud[d].download(), bxs, 0, d, orc.face_shape(g, d)) for d in range(3)]
    phi = B.from_boxed(phid.download(), bxs, 1, g.cell_shape())
    beta = orc.mac_beta_from_density(g, rho, dt)
    uo = [a.copy() for a in u]
    phio, io, ret = orc.mac_project(g, beta, uo, rtol=1e-13)
    assert ret == 0
    div0 = orc.mac_divergence(g, u)
    div1 = orc.mac_divergence(g, un)
    # exact projection: the new divergence is the (tiny) solver residual
    singular = not any(D_ in (bclo[d], bchi[d]) for d in range(3) if not per[d])
    shift = div0.mean() if singular else 0.0
    assert np.max(np.abs(div1 - shift)) <= 1.01 * rtol * np.max(np.abs(div0 - shift))
    tol = 1e-10 if mode == mgx.MODE_FP64 else 1e-7
    assert _rel(phi, phio) < tol
    for d in range(3):
        assert _rel(un[d], uo[d]) < tol
    if mode == mgx.MODE_FP64:
        assert abs(info.iters - io.iters) <= 1
        f = [ctx.array(B.face_field_size(bxs, 0, d)) for d in range(3)]
        P.getFluxes(f[0], f[1], f[2], 0)
        fo = orc.mac_fluxes(g, beta, phio)
        for d in range(3):
            fd = B.faces_from_boxed(f[d].download(), bxs, 0, d, orc.face_shape(g, d))
            assert _rel(fd, fo[d]) < 1e-9
    P.close()


def test_mac_solve_const_beta_decomposition_invariant(ctx, orc):
    from incflo_amd import mgx
    n = (40, 24, 32)
    _say("case mac decomposition invariance")
    g = orc.Geom.make(n, prob_hi=(1.25, 0.75, 1.0))  # isotropic h = 1/32
    rng = np.random.default_rng(11)
    b = rng.standard_normal(g.cell_shape())
    outs = []
    for bxs in (B.chop_domain(n, 64), B.chop_domain(n, (9, 7, 11))):
        P = _mac(ctx, n, bxs, (1, 1, 1), None, None, (1.25, 0.75, 1.0))
        P.updateBeta(0.7)
        rd = ctx.array(B.to_boxed(b, bxs, 0))
        xd = ctx.array(B.field_size(bxs, 0, 1))
        info = P.solve(xd, 0, rd, 0, 1e-12)
        assert info.converged
        outs.append(B.from_boxed(xd.download(), bxs, 0, g.cell_shape()))
        P.close()
    assert _rel(outs[1], outs[0]) < 1e-13
    beta = [np.full(orc.face_shape(g, d), 0.7) for d in range(3)]
    xo, io, _ = orc.mac_solve(g, beta, b, rtol=1e-12)
    assert _rel(outs[0], xo) < 1e-10


def test_mac_error_paths(ctx):
    from incflo_amd import mgx
    _say("case mac error paths")
    bxs = B.chop_domain((16, 16, 16), 8)
    P = mgx.MacProjector(ctx, mgx.make_level((16, 16, 16), bxs, periodic=(0, 1, 1)))
    with pytest.raises(mgx.MgxError) as e:
        P.setDomainBC((mgx.BC_ROBIN, 0, 0), (D_, 0, 0))
    assert e.value.code == mgx.MGX_ERR_UNSUPPORTED
    P.setDomainBC((N_, 0, 0), (D_, 0, 0))
    ud = [ctx.array(B.face_field_size(bxs, 0, d)) for d in range(3)]
    with pytest.raises(mgx.MgxError) as e:
        P.project(ud[0], ud[1], ud[2], 0, 1e-10)
    assert e.value.code == mgx.MGX_ERR_STATE
    P.close()
