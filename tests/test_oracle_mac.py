"""Pins the FP64 CPU oracle's MAC (cell-centred) projection.

Independent constructions: the Fourier symbol of the constant-beta 7-point
operator, a sparse assembly L = D diag(beta) G from separately built
divergence and face-gradient matrices (incl. the Dirichlet/Neumann boundary
faces), a direct sparse solve, the exact-projection property
div(u_new) = residual, and closed forms for the face coefficients.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import Geom, BC_DIRICHLET, BC_NEUMANN, BC_INFLOW

D_, N_ = BC_DIRICHLET, BC_NEUMANN


def _rand(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)


def _betas(orc, g, seed, const=None):
    if const is not None:
        return [np.full(orc.face_shape(g, d), const) for d in range(3)]
    rng = np.random.default_rng(seed)
    out = []
    for d in range(3):
        b = 0.5 + rng.random(orc.face_shape(g, d))
        if g.periodic[d]:  # the periodic face n is face 0
            ax = 2 - d
            idx0 = [slice(None)] * 3; idx0[ax] = 0
            idxn = [slice(None)] * 3; idxn[ax] = -1
            b[tuple(idxn)] = b[tuple(idx0)]
        out.append(b)
    return out


def _dg_matrices(orc, g):
    """Divergence (cells x faces) and face-gradient (faces x cells) matrices."""
    n = list(g.n)
    ncell = n[0] * n[1] * n[2]
    cid = np.arange(ncell).reshape(n[2], n[1], n[0])
    Dblocks, Gblocks = [], []
    for d in range(3):
        fs = orc.face_shape(g, d)
        nf = int(np.prod(fs))
        fid = np.arange(nf).reshape(fs)
        ax = 2 - d
        h = g.dx[d]
        # divergence: + hi face - lo face of each cell
        rows, cols, vals = [], [], []
        lo_faces = np.take(fid, np.arange(n[d]), axis=ax)
        hi_faces = np.take(fid, np.arange(1, n[d] + 1), axis=ax)
        rows += list(cid.ravel()) * 2
        cols += list(hi_faces.ravel()) + list(lo_faces.ravel())
        vals += [1.0 / h] * ncell + [-1.0 / h] * ncell
        Dblocks.append(sp.csr_matrix((vals, (rows, cols)), shape=(ncell, nf)))
        # gradient at faces
        rows, cols, vals = [], [], []
        for f_idx in range(n[d] + 1):
            f = np.take(fid, f_idx, axis=ax).ravel()
            if 0 < f_idx < n[d] or g.periodic[d]:
                L = np.take(cid, (f_idx - 1) % n[d], axis=ax).ravel()
                R = np.take(cid, f_idx % n[d], axis=ax).ravel()
                rows += list(f) * 2; cols += list(R) + list(L)
                vals += [1.0 / h] * len(f) + [-1.0 / h] * len(f)
            elif f_idx == 0 and g.bclo[d] == D_:
                R = np.take(cid, 0, axis=ax).ravel()
                rows += list(f); cols += list(R); vals += [2.0 / h] * len(f)
            elif f_idx == n[d] and g.bchi[d] == D_:
                L = np.take(cid, n[d] - 1, axis=ax).ravel()
                rows += list(f); cols += list(L); vals += [-2.0 / h] * len(f)
        Gblocks.append(sp.csr_matrix((vals, (rows, cols)), shape=(nf, ncell)))
    return Dblocks, Gblocks


def _assembled(orc, g, beta):
    Ds, Gs = _dg_matrices(orc, g)
    return sum(Ds[d] @ sp.diags(beta[d].ravel()) @ Gs[d] for d in range(3))


def test_constant_beta_matches_fft_symbol(orc):
    g = Geom.make((8, 10, 6), prob_hi=(1.0, 2.0, 0.5))
    beta = _betas(orc, g, 0, const=0.3)
    phi = _rand(g.cell_shape(), 1)
    out = orc.mac_apply(g, beta, phi)
    th = [2 * np.pi * np.fft.fftfreq(g.n[d]) for d in range(3)]
    tz, ty, tx = np.meshgrid(th[2], th[1], th[0], indexing="ij")
    lam = -0.3 * sum((2 - 2 * np.cos(t)) / g.dx[d] ** 2 for d, t in enumerate((tx, ty, tz)))
    ref = np.real(np.fft.ifftn(lam * np.fft.fftn(phi)))
    assert np.max(np.abs(out - ref)) < 1e-11 * np.max(np.abs(ref))


@pytest.mark.parametrize("periodic,bclo,bchi", [
    ((1, 1, 1), (0, 0, 0), (0, 0, 0)),
    ((0, 1, 0), (D_, 0, N_), (N_, 0, D_)),
    ((0, 0, 0), (N_, D_, N_), (D_, N_, N_)),
])
def test_variable_beta_matches_dbg_assembly(orc, periodic, bclo, bchi):
    g = Geom.make((6, 5, 7), periodic=periodic, bclo=bclo, bchi=bchi, prob_hi=(1.0, 0.8, 1.4))
    beta = _betas(orc, g, 2)
    phi = _rand(g.cell_shape(), 3)
    A = _assembled(orc, g, beta)
    ref = (A @ phi.ravel()).reshape(g.cell_shape())
    assert np.max(np.abs(orc.mac_apply(g, beta, phi) - ref)) < 1e-11 * np.max(np.abs(ref))
    # operator symmetric and negative semi-definite (self-adjoint D B G)
    assert abs(A - A.T).max() < 1e-9 * abs(A).max()
    assert phi.ravel() @ (A @ phi.ravel()) < 0


def test_fluxes_and_divergence_are_the_assembled_operators(orc):
    g = Geom.make((6, 6, 4), periodic=(0, 1, 1), bclo=(D_, 0, 0), bchi=(N_, 0, 0))
    beta = _betas(orc, g, 4)
    phi = _rand(g.cell_shape(), 5)
    Ds, Gs = _dg_matrices(orc, g)
    f = orc.mac_fluxes(g, beta, phi)
    for d in range(3):
        ref = -(beta[d].ravel() * (Gs[d] @ phi.ravel()))
        assert np.allclose(f[d].ravel(), ref, rtol=1e-13, atol=1e-13)
    u = [_rand(orc.face_shape(g, d), 6 + d) for d in range(3)]
    div = orc.mac_divergence(g, u)
    ref = sum(Ds[d] @ u[d].ravel() for d in range(3)).reshape(g.cell_shape())
    assert np.allclose(div, ref, rtol=1e-13, atol=1e-12)


def test_beta_from_density_closed_form(orc):
    g = Geom.make((4, 3, 5), periodic=(1, 0, 0), bclo=(0, N_, N_), bchi=(0, N_, N_))
    z, y, x = np.meshgrid(np.arange(5), np.arange(3), np.arange(4), indexing="ij")
    rho = 1.0 + x + 10.0 * y + 100.0 * z
    b = orc.mac_beta_from_density(g, rho, 0.1)
    # interior x faces: average of the two neighbours; periodic face 0 pairs cells 3 and 0
    assert b[0][2, 1, 2] == pytest.approx(0.1 / (0.5 * (rho[2, 1, 1] + rho[2, 1, 2])))
    assert b[0][2, 1, 0] == pytest.approx(0.1 / (0.5 * (rho[2, 1, 3] + rho[2, 1, 0])))
    assert np.allclose(b[0][:, :, 0], b[0][:, :, 4])
    # non-periodic boundary face takes the adjacent cell
    assert b[1][1, 0, 1] == pytest.approx(0.1 / rho[1, 0, 1])
    assert b[2][5, 1, 3] == pytest.approx(0.1 / rho[4, 1, 3])
    assert b[2][4, 1, 3] == pytest.approx(0.1 / (0.5 * (rho[3, 1, 3] + rho[4, 1, 3])))
    const = orc.mac_beta_from_density(g, np.full(g.cell_shape(), 2.0), 0.5)
    assert all(np.allclose(c, 0.25) for c in const)


def test_beta_coarsen_is_face_average(orc):
    gf = Geom.make((8, 8, 8))
    bf = [_rand(orc.face_shape(gf, d), 10 + d) for d in range(3)]
    bc = orc.mac_beta_coarsen(gf, bf[0], 0)
    # coarse x-face (I,J,K) = mean of fine x-faces (2I, 2J+{0,1}, 2K+{0,1})
    I, J, K = 2, 1, 3
    ref = bf[0][2 * K:2 * K + 2, 2 * J:2 * J + 2, 2 * I].mean()
    assert bc[K, J, I] == pytest.approx(ref, rel=1e-14)


@pytest.mark.parametrize("periodic,bclo,bchi", [
    ((0, 0, 0), (D_, N_, N_), (D_, N_, N_)),
    ((0, 1, 1), (N_, 0, 0), (D_, 0, 0)),
])
def test_mac_solve_matches_sparse_direct(orc, periodic, bclo, bchi):
    g = Geom.make((16, 8, 8), periodic=periodic, bclo=bclo, bchi=bchi, prob_hi=(2.0, 1.0, 1.0))
    beta = _betas(orc, g, 12)
    b = _rand(g.cell_shape(), 13)
    x, info, ret = orc.mac_solve(g, beta, b, rtol=1e-12)
    assert ret == 0 and info.converged
    xd = spla.spsolve(_assembled(orc, g, beta).tocsc(), b.ravel())
    assert np.max(np.abs(x.ravel() - xd)) < 1e-9 * np.max(np.abs(xd))


def test_mac_periodic_constant_beta_matches_fft_and_contracts(orc):
    g = Geom.make((32, 32, 32))
    beta = _betas(orc, g, 0, const=1.5)
    b = _rand(g.cell_shape(), 14); b -= b.mean()
    x, info, ret = orc.mac_solve(g, beta, b, rtol=1e-12)
    assert ret == 0
    th = [2 * np.pi * np.fft.fftfreq(32)] * 3
    tz, ty, tx = np.meshgrid(*th, indexing="ij")
    lam = -1.5 * sum((2 - 2 * np.cos(t)) / g.dx[0] ** 2 for t in (tx, ty, tz)); lam[0, 0, 0] = 1
    xh = np.fft.fftn(b) / lam; xh[0, 0, 0] = 0
    xf = np.real(np.fft.ifftn(xh))
    assert np.max(np.abs(x - xf)) < 1e-9 * np.max(np.abs(xf))
    assert (info.resnorm / info.resnorm0) ** (1.0 / info.iters) < 0.35


def test_mac_projection_is_exact(orc):
    """After the projection the discrete divergence equals the solver residual."""
    g = Geom.make((24, 16, 16), periodic=(0, 1, 1), bclo=(N_, 0, 0), bchi=(D_, 0, 0),
                  prob_hi=(1.5, 1.0, 1.0))
    rng = np.random.default_rng(15)
    rho = 1.0 + 0.5 * np.tanh(rng.standard_normal(g.cell_shape()))
    beta = orc.mac_beta_from_density(g, rho, 0.01)
    u = [rng.standard_normal(orc.face_shape(g, d)) for d in range(3)]
    u[0][:, :, 0] = 1.0  # inflow face value is kept (Neumann)
    div0 = orc.mac_divergence(g, u)
    phi, info, ret = orc.mac_project(g, beta, u, rtol=1e-12)
    assert ret == 0
    div1 = orc.mac_divergence(g, u)
    assert np.max(np.abs(div1)) <= 1.01e-12 * np.max(np.abs(div0))
    assert np.all(u[0][:, :, 0] == 1.0)
