"""Pins the FP64 CPU oracle's nodal operator, D, G, transfer operators and MLMG.

The checks do not retype the oracle's formulas: each one uses an independent
construction (FFT symbol of the 1-D stiffness/mass product, a Kronecker-product
sparse assembly, quadratic exactness, adjointness, a direct sparse solve, the
Galerkin identity) chosen so that a dropped term, a wrong sign or index or a
wrong h-scaling fails at least one of them.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import Geom, BC_DIRICHLET, BC_NEUMANN, BC_INFLOW


def _rand(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)


def _symbol(n, dx, sigma):
    """Closed-form Fourier symbol of the constant-sigma Q1 nodal operator."""
    th = [2 * np.pi * np.fft.fftfreq(n[d]) for d in range(3)]
    tz, ty, tx = np.meshgrid(th[2], th[1], th[0], indexing="ij")
    t = [tx, ty, tz]
    lam = np.zeros_like(tx)
    for d in range(3):
        term = (2 - 2 * np.cos(t[d])) / dx[d] ** 2
        for e in range(3):
            if e != d:
                term = term * (2 + np.cos(t[e])) / 3
        lam += term
    return -sigma * lam


def _kron_matrix(g, sigma):
    """Global matrix of L assembled from 1-D element matrices by Kronecker products."""
    K1 = np.array([[1.0, -1.0], [-1.0, 1.0]])
    M1 = np.array([[2.0, 1.0], [1.0, 2.0]]) / 6.0
    dx = g.dx
    # local index m = mx + 2 my + 4 mz  ->  kron(z, y, x)
    Ek = (np.kron(M1, np.kron(M1, K1)) / dx[0] ** 2 + np.kron(M1, np.kron(K1, M1)) / dx[1] ** 2
          + np.kron(K1, np.kron(M1, M1)) / dx[2] ** 2)
    nx, ny, nz = g.nodes()
    rows, cols, vals = [], [], []
    for k in range(g.n[2]):
        for j in range(g.n[1]):
            for i in range(g.n[0]):
                s = sigma[k, j, i]
                ids = []
                for m in range(8):
                    a, b, c = i + (m & 1), j + ((m >> 1) & 1), k + ((m >> 2) & 1)
                    if g.periodic[0]: a %= nx
                    if g.periodic[1]: b %= ny
                    if g.periodic[2]: c %= nz
                    ids.append(a + nx * (b + ny * c))
                for p in range(8):
                    for m in range(8):
                        rows.append(ids[p]); cols.append(ids[m]); vals.append(-s * Ek[p, m])
    N = nx * ny * nz
    return sp.csr_matrix((vals, (rows, cols)), shape=(N, N))


def test_constant_sigma_matches_fft_symbol(orc):
    g = Geom.make((8, 6, 10), prob_hi=(1.0, 0.75, 2.5))
    sigma = np.full(g.cell_shape(), 0.7)
    phi = _rand(g.node_shape(), 1)
    out = orc.nodal_apply(g, sigma, phi)
    ref = np.real(np.fft.ifftn(_symbol(g.n, g.dx, 0.7) * np.fft.fftn(phi)))
    assert np.max(np.abs(out - ref)) < 1e-10 * np.max(np.abs(ref))


def test_isotropic_stencil_weights(orc):
    """Centre -8/3, faces 0, edges 1/6, corners 1/12 (sigma=1, h=1)."""
    g = Geom.make((6, 6, 6), prob_hi=(6.0, 6.0, 6.0))
    sigma = np.ones(g.cell_shape())
    phi = np.zeros(g.node_shape()); phi[3, 3, 3] = 1.0
    out = orc.nodal_apply(g, sigma, phi)
    assert out[3, 3, 3] == pytest.approx(-8.0 / 3.0, abs=1e-14)
    assert out[3, 3, 4] == pytest.approx(0.0, abs=1e-14)
    assert out[3, 4, 4] == pytest.approx(1.0 / 6.0, abs=1e-14)
    assert out[4, 4, 4] == pytest.approx(1.0 / 12.0, abs=1e-14)


@pytest.mark.parametrize("periodic", [(1, 1, 1), (0, 1, 0), (0, 0, 0)])
def test_variable_sigma_matches_kron_assembly(orc, periodic):
    g = Geom.make((5, 4, 6), periodic=periodic, bclo=(BC_NEUMANN,) * 3, bchi=(BC_NEUMANN,) * 3,
                  prob_hi=(1.0, 0.8, 1.5))
    sigma = 0.5 + np.random.default_rng(3).random(g.cell_shape())
    phi = _rand(g.node_shape(), 4)
    A = _kron_matrix(g, sigma)
    ref = (A @ phi.ravel()).reshape(g.node_shape())
    out = orc.nodal_apply(g, sigma, phi)
    assert np.max(np.abs(out - ref)) < 1e-11 * np.max(np.abs(ref))
    # diagonal
    assert np.allclose(orc.nodal_diag(g, sigma).ravel(), A.diagonal(), rtol=1e-13, atol=0)


def test_operator_invariants(orc):
    g = Geom.make((6, 5, 7), periodic=(0, 0, 1), bclo=(BC_NEUMANN,) * 3, bchi=(BC_NEUMANN,) * 3)
    sigma = 0.5 + np.random.default_rng(5).random(g.cell_shape())
    a, b = _rand(g.node_shape(), 6), _rand(g.node_shape(), 7)
    La, Lb = orc.nodal_apply(g, sigma, a), orc.nodal_apply(g, sigma, b)
    assert abs(np.sum(La * b) - np.sum(a * Lb)) < 1e-11 * np.sum(np.abs(La * b))  # symmetric
    assert np.sum(La * a) < 0  # negative definite on non-constants
    assert np.max(np.abs(orc.nodal_apply(g, sigma, np.ones(g.node_shape())))) < 1e-12  # L 1 = 0


def test_quadratic_exactness(orc):
    """L(x^2) = 2 sigma and L(xy) = 0 at interior nodes (constant sigma)."""
    g = Geom.make((8, 8, 8), periodic=(0, 0, 0), bclo=(BC_NEUMANN,) * 3, bchi=(BC_NEUMANN,) * 3,
                  prob_hi=(1.0, 2.0, 0.5))
    sigma = np.full(g.cell_shape(), 1.3)
    nx, ny, nz = g.nodes()
    z, y, x = np.meshgrid(np.arange(nz) * g.dx[2], np.arange(ny) * g.dx[1],
                          np.arange(nx) * g.dx[0], indexing="ij")
    inner = (slice(1, -1),) * 3
    assert np.allclose(orc.nodal_apply(g, sigma, x ** 2)[inner], 2 * 1.3, rtol=1e-10)
    assert np.allclose(orc.nodal_apply(g, sigma, y ** 2 + z ** 2)[inner], 4 * 1.3, rtol=1e-10)
    assert np.max(np.abs(orc.nodal_apply(g, sigma, x * y)[inner])) < 1e-9


def test_div_grad_adjoint_and_linear_exactness(orc):
    g = Geom.make((6, 8, 4), prob_hi=(1.5, 1.0, 0.5))
    ug = np.zeros(g.ghost_vec_shape())
    ug[:, 1:-1, 1:-1, 1:-1] = _rand((3,) + g.cell_shape(), 8)
    orc.fill_vel_ghosts(g, ug)
    phi = _rand(g.node_shape(), 9)
    Du = orc.nodal_divergence(g, ug)
    Gp = orc.nodal_gradient(g, phi)
    lhs = np.sum(Du * phi)
    rhs = -np.sum(ug[:, 1:-1, 1:-1, 1:-1] * Gp)
    assert abs(lhs - rhs) < 1e-11 * np.sum(np.abs(Du * phi))
    # D of a linear field (non-periodic, Neumann walls): interior nodes give the exact divergence
    g2 = Geom.make((6, 6, 6), periodic=(0, 0, 0), bclo=(BC_NEUMANN,) * 3, bchi=(BC_NEUMANN,) * 3)
    ug2 = np.zeros(g2.ghost_vec_shape())
    zc, yc, xc = np.meshgrid(*(((np.arange(-1, 7) + 0.5) * g2.dx[d]) for d in (2, 1, 0)),
                             indexing="ij")
    ug2[0], ug2[1], ug2[2] = 2 * xc, -3 * yc, 0.5 * zc
    d2 = orc.nodal_divergence(g2, ug2)
    assert np.allclose(d2, 2 - 3 + 0.5, atol=1e-12)
    # G of a linear phi is exact
    nz_, ny_, nx_ = g2.node_shape()
    zn, yn, xn = np.meshgrid(np.arange(nz_) * g2.dx[2], np.arange(ny_) * g2.dx[1],
                             np.arange(nx_) * g2.dx[0], indexing="ij")
    gp2 = orc.nodal_gradient(g2, 1.5 * xn - 2 * yn + 4 * zn)
    assert np.allclose(gp2[0], 1.5) and np.allclose(gp2[1], -2) and np.allclose(gp2[2], 4)


def test_inflow_ghost_rule(orc):
    g = Geom.make((4, 4, 4), periodic=(0, 1, 1), bclo=(BC_INFLOW, 0, 0), bchi=(BC_DIRICHLET, 0, 0))
    ug = np.zeros(g.ghost_vec_shape())
    iv = np.zeros((6, 3)); iv[0] = [1.0, 0.0, 0.0]; iv[1] = [5.0, 5.0, 5.0]
    orc.fill_vel_ghosts(g, ug, iv)
    assert np.all(ug[0, 1:-1, 1:-1, 0] == 1.0)   # x-lo inflow face
    assert np.all(ug[0, :, :, -1] == 0.0)        # x-hi outflow: not an inflow face -> 0
    assert np.all(ug[1:, 1:-1, 1:-1, 0] == 0.0)


def _prolong_matrix(orc, gf):
    gc = gf.coarsen()
    Nc = int(np.prod(gc.node_shape()))
    cols = []
    for c in range(Nc):
        e = np.zeros(Nc); e[c] = 1.0
        cols.append(orc.nodal_prolong_add(gf, e.reshape(gc.node_shape()),
                                          np.zeros(gf.node_shape())).ravel())
    return np.array(cols).T


@pytest.mark.parametrize("periodic", [(1, 1, 1), (0, 1, 0)])
def test_restriction_is_scaled_prolongation_transpose(orc, periodic):
    gf = Geom.make((8, 8, 8), periodic=periodic, bclo=(BC_NEUMANN,) * 3, bchi=(BC_NEUMANN,) * 3)
    P = _prolong_matrix(orc, gf)
    # trilinear: rows sum to 1, coarse node values are copied
    assert np.allclose(P.sum(axis=1), 1.0)
    r = _rand(gf.node_shape(), 10)
    Rr = orc.nodal_restrict(gf, r).ravel()
    assert np.allclose(Rr, P.T @ r.ravel() / 8.0, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("periodic", [(1, 1, 1), (0, 0, 1)])
def test_galerkin_identity_constant_sigma(orc, periodic):
    """L_2h = P^T L_h P / 8 for nested Q1 spaces with constant sigma."""
    gf = Geom.make((8, 8, 8), periodic=periodic, bclo=(BC_NEUMANN,) * 3, bchi=(BC_NEUMANN,) * 3,
                   prob_hi=(1.0, 1.0, 2.0))
    gc = gf.coarsen()
    sf = np.full(gf.cell_shape(), 0.9)
    sc = orc.sigma_coarsen(gf, sf)
    assert np.allclose(sc, 0.9)
    P = _prolong_matrix(orc, gf)
    xc = _rand(gc.node_shape(), 11)
    Lc = orc.nodal_apply(gc, sc, xc).ravel()
    Lf = orc.nodal_apply(gf, sf, (P @ xc.ravel()).reshape(gf.node_shape())).ravel()
    assert np.allclose(Lc, P.T @ Lf / 8.0, rtol=1e-11, atol=1e-11)


def test_bicgstab_matches_direct_solve(orc):
    g = Geom.make((4, 4, 4), periodic=(0, 0, 0), bclo=(BC_DIRICHLET, BC_NEUMANN, BC_NEUMANN),
                  bchi=(BC_NEUMANN,) * 3)
    sigma = 0.5 + np.random.default_rng(12).random(g.cell_shape())
    b = _rand(g.node_shape(), 13); b[:, :, 0] = 0.0
    x, its, ret = orc.nodal_bicgstab(g, sigma, b, rtol=1e-13, maxiter=500)
    A = _kron_matrix(g, sigma).toarray()
    free = np.ones(g.node_shape(), bool); free[:, :, 0] = False
    f = free.ravel()
    xd = np.zeros(f.size)
    xd[f] = np.linalg.solve(A[np.ix_(f, f)], b.ravel()[f])
    assert ret == 0
    assert np.max(np.abs(x.ravel() - xd)) < 1e-10 * np.max(np.abs(xd))


def test_mlmg_periodic_constant_sigma_matches_fft(orc):
    g = Geom.make((32, 32, 32))
    sigma = np.full(g.cell_shape(), 2.0)
    b = _rand(g.node_shape(), 14); b -= b.mean()
    x, info, ret = orc.nodal_solve(g, sigma, b, rtol=1e-12)
    assert ret == 0 and info.converged
    lam = _symbol(g.n, g.dx, 2.0); lam[0, 0, 0] = 1.0
    xh = np.fft.fftn(b) / lam; xh[0, 0, 0] = 0.0
    xf = np.real(np.fft.ifftn(xh))
    assert np.max(np.abs(x - xf)) < 1e-9 * np.max(np.abs(xf))
    assert info.iters < 20


@pytest.mark.parametrize("periodic,bclo,bchi", [
    ((0, 0, 0), (BC_DIRICHLET, BC_NEUMANN, BC_NEUMANN), (BC_NEUMANN, BC_NEUMANN, BC_DIRICHLET)),
    ((0, 1, 1), (BC_INFLOW, 0, 0), (BC_DIRICHLET, 0, 0)),
])
def test_mlmg_variable_sigma_matches_sparse_direct(orc, periodic, bclo, bchi):
    g = Geom.make((16, 8, 8), periodic=periodic, bclo=bclo, bchi=bchi, prob_hi=(2.0, 1.0, 1.0))
    sigma = 0.5 + np.random.default_rng(15).random(g.cell_shape())
    b = _rand(g.node_shape(), 16)
    x, info, ret = orc.nodal_solve(g, sigma, b, rtol=1e-12)
    assert ret == 0
    A = _kron_matrix(g, sigma)
    dmask = np.zeros(g.node_shape(), bool)
    nx, ny, nz = g.nodes()
    for d, (lo, hi) in enumerate(zip(bclo, bchi)):
        if periodic[d]:
            continue
        sl_lo = [slice(None)] * 3; sl_lo[2 - d] = 0
        sl_hi = [slice(None)] * 3; sl_hi[2 - d] = -1
        if lo == BC_DIRICHLET: dmask[tuple(sl_lo)] = True
        if hi == BC_DIRICHLET: dmask[tuple(sl_hi)] = True
    f = ~dmask.ravel()
    xd = np.zeros(f.size)
    xd[f] = spla.spsolve(A[f][:, f].tocsc(), b.ravel()[f])
    assert np.max(np.abs(x.ravel() - xd)) < 1e-9 * np.max(np.abs(xd))
    r = orc.nodal_residual(g, sigma, np.where(dmask, 0, b), x)
    assert np.max(np.abs(r)) <= 1e-12 * np.max(np.abs(np.where(dmask, 0, b))) * 1.0001


def test_mlmg_vcycle_contracts_and_is_smoother_independent(orc):
    g = Geom.make((32, 32, 32))
    sigma = 1.0 + 0.5 * np.tanh(_rand(g.cell_shape(), 17))
    b = _rand(g.node_shape(), 18); b -= b.mean()
    x1, i1, _ = orc.nodal_solve(g, sigma, b, rtol=1e-11)
    x2, i2, _ = orc.nodal_solve(g, sigma, b, rtol=1e-11,
                                opts=orc.default_opts(nu1=3, nu2=1, omega=2.0 / 3.0))
    assert i1.converged and i2.converged
    # average contraction per V-cycle
    rho = (i1.resnorm / i1.resnorm0) ** (1.0 / i1.iters)
    assert rho < 0.2, rho
    # same discrete solution regardless of smoother settings (mean-free)
    assert np.max(np.abs(x1 - x2)) < 1e-8 * np.max(np.abs(x1))


@pytest.mark.parametrize("hr", [(1, 1, 1), (1 / 40, 1 / 24, 1 / 32), (1, 2, 1), (1, 3, 3)])
def test_jacobi_omega_scaling_matches_symbol_peak(orc, hr):
    """The per-level omega uses the peak of the Jacobi-preconditioned symbol; check
    the closed form against the symbol evaluated on a fine Fourier grid."""
    import ctypes
    n = (16, 16, 16)
    g = Geom.make(n, prob_hi=tuple(n[d] * hr[d] for d in range(3)))
    lam = -_symbol(n, g.dx, 1.0)
    diag = 8.0 / 9.0 * sum(1.0 / g.dx[d] ** 2 for d in range(3))
    peak = lam.max() / diag
    orc.lib().orc_nodal_omega.restype = ctypes.c_double
    w = orc.lib().orc_nodal_omega(ctypes.byref(g), ctypes.c_double(8.0 / 9.0))
    assert w == pytest.approx(8.0 / 9.0 * min(1.0, 1.5 / peak), rel=1e-12)
    assert abs(1.0 - w * peak) < 1.0


def test_mlmg_converges_on_anisotropic_cells(orc):
    g = Geom.make((40, 24, 32))  # h = 1/40, 1/24, 1/32
    sigma = 1.0 + 0.5 * np.tanh(_rand(g.cell_shape(), 20))
    b = _rand(g.node_shape(), 21); b -= b.mean()
    x, info, ret = orc.nodal_solve(g, sigma, b, rtol=1e-11)
    assert ret == 0 and info.converged
    assert (info.resnorm / info.resnorm0) ** (1.0 / info.iters) < 0.4


def test_projection_reduces_divergence_to_approximate_level(orc):
    """Approximate projection: L phi = D u holds to solver precision, D(u_new) is only small."""
    g = Geom.make((16, 16, 16))
    ug = np.zeros(g.ghost_vec_shape())
    ug[:, 1:-1, 1:-1, 1:-1] = _rand((3,) + g.cell_shape(), 19)
    orc.fill_vel_ghosts(g, ug)
    sigma = np.full(g.cell_shape(), 0.01)
    div0 = orc.nodal_divergence(g, ug)
    phi, gphi, info, ret = orc.nodal_project(g, sigma, ug, rtol=1e-12)
    assert ret == 0
    r = orc.nodal_residual(g, sigma, div0 - div0.mean(), phi)
    assert np.max(np.abs(r)) <= 1e-12 * np.max(np.abs(div0)) * 1.01
    orc.fill_vel_ghosts(g, ug)
    div1 = orc.nodal_divergence(g, ug)
    assert np.linalg.norm(div1) < 0.6 * np.linalg.norm(div0)
