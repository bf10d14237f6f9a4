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
    P.setDomainBC((N_, 0, 0), (D_, 0, 0))
    ud = [ctx.array(B.face_field_size(bxs, 0, d)) for d in range(3)]
    with pytest.raises(mgx.MgxError) as e:
        P.project(ud[0], ud[1], ud[2], 0, 1e-10)
    assert e.value.code == mgx.MGX_ERR_STATE
    # a Robin side needs setLevelBC's coefficients before a solve
    P.setDomainBC((mgx.BC_ROBIN, 0, 0), (D_, 0, 0))
    P.updateBeta(1.0)
    with pytest.raises(mgx.MgxError) as e:
        P.project(ud[0], ud[1], ud[2], 0, 1e-10)
    assert e.value.code == mgx.MGX_ERR_STATE
    # b < 0 is refused
    A = np.zeros((1, 18, 18, 18))
    A[0, :, :, 0] = 1.0
    Bv = np.full((1, 18, 18, 18), -1.0)
    bo = [ctx.array(B.set_ghosts_from(np.zeros(B.field_size(bxs, 1)), bxs, 1, X, ncomp=1))
          for X in (A, Bv, np.zeros_like(A))]
    with pytest.raises(mgx.MgxError) as e:
        P.setLevelBC(bo[0], bo[1], bo[2], 1)
    assert e.value.code == mgx.MGX_ERR_ARG
    P.close()


def _robin_boxed(ctx, g, bxs, rob):
    """Robin planes -> three boxed cell fields (ng 1) holding them in the
    ghost cells outside the faces (incflo's make_robinBC_MFs layout)."""
    nz, ny, nx = g.cell_shape()
    out = []
    for q in range(3):
        X = np.zeros((1, nz + 2, ny + 2, nx + 2))
        for (d, s), planes in rob.items():
            v = planes[q]
            gi = 0 if s == 0 else g.n[d] + 1
            if d == 0:
                X[0, 1:-1, 1:-1, gi] = v
            elif d == 1:
                X[0, 1:-1, gi, 1:-1] = v
            else:
                X[0, gi, 1:-1, 1:-1] = v
        out.append(ctx.array(B.set_ghosts_from(np.zeros(B.field_size(bxs, 1)), bxs, 1, X,
                                               ncomp=1)))
    return out


@pytest.mark.parametrize("mode", [0, 1])
def test_mac_robin_mixed_face_projection(ctx, orc, mode):
    """incflo's mixed face in the MAC projection: Robin with a = 1, b = 0 on
    the outflow half and a = 0, b = 1 on the inflow half, f = 0
    (src/prob/prob_bc.cpp:97-197) against the oracle (oracle/mac_robin.py)."""
    from incflo_amd import mgx
    from oracle import mac_robin as R
    _say(f"case mac robin mixed face mode={mode}")
    n, per, ph = (32, 16, 16), (0, 0, 1), (2.0, 1.0, 1.0)
    bclo, bchi = (N_, N_, 0), (mgx.BC_ROBIN, N_, 0)
    g = orc.Geom.make(n, periodic=per, bclo=bclo, bchi=bchi, prob_hi=ph)
    rng = np.random.default_rng(21)
    rho = 1.0 + 0.5 * np.tanh(rng.standard_normal(g.cell_shape()))
    u = _periodic_consistent(orc, g, [rng.standard_normal(orc.face_shape(g, d)) for d in range(3)])
    dt = 0.01
    shp = R.plane_shape(g, 0)
    a = np.zeros(shp)
    a[:, : shp[1] // 2] = 1.0
    rob = {(0, 1): (a, 1.0 - a, np.zeros(shp))}
    bxs = B.chop_domain(n, (16, 8, 8))
    P = _mac(ctx, n, bxs, per, bclo, bchi, ph)
    P.set_beta_from_density(ctx.array(B.to_boxed(rho, bxs, 1, periodic=per)), 1, dt)
    ra, rb, rf = _robin_boxed(ctx, g, bxs, rob)
    P.setLevelBC(ra, rb, rf, 1)
    ud = [ctx.array(B.faces_to_boxed(u[d], bxs, 0, d)) for d in range(3)]
    phid = ctx.array(B.field_size(bxs, 0, 1))
    rtol = 1e-12 if mode == mgx.MODE_FP64 else 1e-10
    info = P.project(ud[0], ud[1], ud[2], 0, rtol, 0.0, phi=phid, phi_ng=0, mode=mode)
    assert info.converged
    un = [B.faces_from_boxed(ud[d].download(), bxs, 0, d, orc.face_shape(g, d)) for d in range(3)]
    phi = B.from_boxed(phid.download(), bxs, 0, g.cell_shape())
    beta = orc.mac_beta_from_density(g, rho, dt)
    uo, phio = R.robin_project(g, beta, rob, u)
    tol = 1e-9 if mode == mgx.MODE_FP64 else 1e-7
    assert _rel(phi, phio) < tol
    for d in range(3):
        assert _rel(un[d], uo[d]) < tol, d
    top = a == 0
    assert np.array_equal(un[0][:, :, -1][top], u[0][:, :, -1][top])
    f = [ctx.array(B.face_field_size(bxs, 0, d)) for d in range(3)]
    P.getFluxes(f[0], f[1], f[2], 0)
    Fo = R.robin_fluxes(g, beta, rob, phio)
    for d in range(3):
        fd = B.faces_from_boxed(f[d].download(), bxs, 0, d, orc.face_shape(g, d))
        assert _rel(fd, -Fo[d]) < tol, d
    P.close()


def test_mac_robin_general_solve(ctx, orc):
    """MLMG::solve on the MAC operator with general Robin data (a, b > 0,
    f != 0) on two sides, Dirichlet nowhere, against the oracle's direct
    solve; a = 1, b = 0, f = 0 reproduces the Dirichlet side exactly."""
    from incflo_amd import mgx
    from oracle import mac_robin as R
    _say("case mac robin general solve")
    n, per, ph = (24, 16, 32), (0, 0, 0), (1.5, 1.0, 2.0)
    bclo, bchi = (mgx.BC_ROBIN, N_, N_), (N_, mgx.BC_ROBIN, N_)
    g = orc.Geom.make(n, periodic=per, bclo=bclo, bchi=bchi, prob_hi=ph)
    rng = np.random.default_rng(22)
    beta = [0.5 + rng.random(orc.face_shape(g, d)) for d in range(3)]
    rob = {}
    for (d, s) in [(0, 0), (1, 1)]:
        shp = R.plane_shape(g, d)
        rob[(d, s)] = (rng.random(shp), 0.05 + 0.1 * rng.random(shp), rng.standard_normal(shp))
    rhs = rng.standard_normal(g.cell_shape())
    bxs = B.chop_domain(n, (8, 16, 16))
    P = _mac(ctx, n, bxs, per, bclo, bchi, ph)
    P.set_beta(*[ctx.array(B.faces_to_boxed(beta[d], bxs, 0, d)) for d in range(3)], 0)
    ra, rb, rf = _robin_boxed(ctx, g, bxs, rob)
    P.setLevelBC(ra, rb, rf, 1)
    rd = ctx.array(B.to_boxed(rhs, bxs, 0))
    xd = ctx.array(B.field_size(bxs, 0, 1))
    info = P.solve(xd, 0, rd, 0, 1e-12)
    assert info.converged
    x = B.from_boxed(xd.download(), bxs, 0, g.cell_shape())
    xo = R.robin_solve(g, beta, rob, rhs)
    assert _rel(x, xo) < 1e-9
    P.close()
    # Dirichlet limit: identical to a Dirichlet x-lo side
    rob1 = {(d, s): (np.ones_like(v[0]), np.zeros_like(v[0]), np.zeros_like(v[0]))
            for (d, s), v in rob.items()}
    outs = []
    for lo, hi in [(bclo, bchi), ((D_, N_, N_), (N_, D_, N_))]:
        P = _mac(ctx, n, bxs, per, lo, hi, ph)
        P.set_beta(*[ctx.array(B.faces_to_boxed(beta[d], bxs, 0, d)) for d in range(3)], 0)
        if lo[0] == mgx.BC_ROBIN:
            P.setLevelBC(*_robin_boxed(ctx, g, bxs, rob1), 1)
        xd = ctx.array(B.field_size(bxs, 0, 1))
        info = P.solve(xd, 0, rd, 0, 1e-12)
        assert info.converged
        outs.append(xd.download())
        P.close()
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("mode", [0, 1])
def test_mac_graph_replay_matches_plain_launches(ctx, orc, mode):
    """The cell-centred V-cycle replayed from HIP graphs (captured on the first
    cycle, replayed afterwards) gives bit for bit the face velocities and phi
    of plain launches: three projections each way on one projector, FP64 and
    mixed, in the bounded channel layout (inflow and Dirichlet faces)."""
    n, per, bclo, bchi, ph = CASES["channel"]
    _say(f"case mac graphs vs plain launches mode={mode}")
    g = orc.Geom.make(n, periodic=per, bclo=bclo, bchi=bchi, prob_hi=ph)
    rng = np.random.default_rng(9)
    rho = 1.0 + 0.5 * np.tanh(rng.standard_normal(g.cell_shape()))
    u = _periodic_consistent(orc, g, [rng.standard_normal(orc.face_shape(g, d)) for d in range(3)])
    u[0][:, :, 0] = 1.0
    bxs = B.chop_domain(n, (16, 8, 12))
    P = _mac(ctx, n, bxs, per, bclo, bchi, ph)
    P.set_beta_from_density(ctx.array(B.to_boxed(rho, bxs, 1, periodic=per)), 1, 0.01)
    runs = {}
    for graphs in (True, False):
        P.set_graphs(graphs)
        out = []
        for rep in range(3):
            ud = [ctx.array(B.faces_to_boxed(u[d], bxs, 0, d)) for d in range(3)]
            phid = ctx.array(B.field_size(bxs, 1, 1))
            info = P.project(ud[0], ud[1], ud[2], 0, 1e-11, 0.0, phi=phid, phi_ng=1, mode=mode)
            assert info.converged
            ctx.sync()
            out.append((info.iters, phid.download(), [a.download() for a in ud]))
        runs[graphs] = out
    P.close()
    ref = runs[False][0]
    for it, phi, us in runs[True] + runs[False][1:]:
        assert it == ref[0]
        assert np.array_equal(phi, ref[1])
        for d in range(3):
            assert np.array_equal(us[d], ref[2][d])


@pytest.mark.parametrize("mode", [0, 1])
def test_mac_coarse_cycle_matches_separate_launches(ctx, orc, mode):
    """The coarse levels of the cell-centred V-cycle run in one workgroup
    (levels of at most mac_coarse_cells cells) give bit for bit the projection
    of one launch per pass (mac_coarse_cells 0): same per-cell arithmetic, a
    barrier between passes.  Bounded channel (inflow, Dirichlet) and the
    periodic case, FP64 and mixed."""
    from incflo_amd import mgx
    base = mgx.tuning_get()
    lines = [ln for ln in base.splitlines() if not ln.startswith("mac_coarse_cells")]
    try:
        for case in ("channel", "periodic"):
            n, per, bclo, bchi, ph = CASES[case]
            _say(f"case mac coarse cycle {case} mode={mode}")
            g = orc.Geom.make(n, periodic=per, bclo=bclo, bchi=bchi, prob_hi=ph)
            rng = np.random.default_rng(13)
            rho = 1.0 + 0.5 * np.tanh(rng.standard_normal(g.cell_shape()))
            u = _periodic_consistent(orc, g, [rng.standard_normal(orc.face_shape(g, d)) for d in range(3)])
            if not per[0]:
                u[0][:, :, 0] = 1.0
            bxs = B.chop_domain(n, (16, 8, 12))
            out = {}
            for cells in (4096, 0):
                mgx.tuning_set("\n".join(lines + [f"mac_coarse_cells {cells}"]) + "\n")
                P = _mac(ctx, n, bxs, per, bclo, bchi, ph)
                P.set_beta_from_density(ctx.array(B.to_boxed(rho, bxs, 1, periodic=per)), 1, 0.01)
                ud = [ctx.array(B.faces_to_boxed(u[d], bxs, 0, d)) for d in range(3)]
                phid = ctx.array(B.field_size(bxs, 1, 1))
                info = P.project(ud[0], ud[1], ud[2], 0, 1e-11, 0.0, phi=phid, phi_ng=1, mode=mode)
                assert info.converged
                out[cells] = (info.iters, phid.download(), [a.download() for a in ud])
                P.close()
            assert out[4096][0] == out[0][0]
            assert np.array_equal(out[4096][1], out[0][1])
            for d in range(3):
                assert np.array_equal(out[4096][2][d], out[0][2][d])
    finally:
        mgx.tuning_set(base)
