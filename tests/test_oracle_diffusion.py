"""Pins of the diffusion oracle (oracle/diffusion.py): the cell-centred
viscous operator and the Helmholtz solve behind incflo's implicit viscous
step (DiffusionScalarOp::diffuse_vel_components,
src/diffusion/DiffusionScalarOp.cpp:332-480).

Independent checks: the FFT symbol on periodic grids (constant coefficients),
a sparse direct solve, symmetry and row sums, second-order convergence to a
manufactured solution with Dirichlet and Neumann faces, and the face
averaging of eta against its definition on linear fields."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import BC_DIRICHLET, BC_NEUMANN, Geom
from oracle import diffusion as D


def _fft_helmholtz(g, a, b, beta, rhs):
    k = [2 * np.pi * np.fft.fftfreq(g.n[d]) for d in range(3)]
    sym = np.zeros(g.cell_shape())
    for d in range(3):
        shape = [1, 1, 1]
        shape[2 - d] = g.n[d]
        sym = sym + (2 * np.cos(k[d]) - 2).reshape(shape) / g.dx[d] ** 2
    return np.real(np.fft.ifftn(np.fft.fftn(rhs) / (a - beta * b * sym)))


def test_periodic_helmholtz_equals_fft_solution():
    g = Geom.make((12, 10, 8), prob_hi=(1.2, 1.0, 0.8))
    rng = np.random.default_rng(1)
    rhs = rng.standard_normal(g.cell_shape())
    a, b, beta = 1.3, 0.7, 0.05
    bf = D.eta_faces(np.full(g.cell_shape(), b), g.periodic)
    u = D.helmholtz_solve(g, np.full(g.cell_shape(), a), bf, beta, rhs)
    ref = _fft_helmholtz(g, a, b, beta, rhs)
    assert np.max(np.abs(u - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_operator_symmetric_row_sums_and_direct_solve():
    g = Geom.make((6, 5, 4), periodic=(0, 1, 0), bclo=(BC_DIRICHLET, 0, BC_NEUMANN),
                  bchi=(BC_NEUMANN, 0, BC_DIRICHLET))
    rng = np.random.default_rng(2)
    eta = 0.5 + rng.random(g.cell_shape())
    bf = D.eta_faces(eta, g.periodic)
    A, k = D.operator(g, bf)
    assert abs(A - A.T).max() == 0.0
    # rows of cells off Dirichlet faces sum to zero (constants in the kernel)
    rs = np.asarray(A.sum(axis=1)).reshape(g.cell_shape())
    inner = rs.copy()
    inner[:, :, 0] = 0.0   # x lo Dirichlet
    inner[-1, :, :] = 0.0  # z hi Dirichlet
    assert np.max(np.abs(inner)) <= 1e-12
    assert np.all(rs[:, :, 0] < 0) and np.all(rs[-1, :, :] < 0)
    # CG answer of the Helmholtz system equals a sparse direct solve
    a = 1.0 + rng.random(g.cell_shape())
    rhs = rng.standard_normal(g.cell_shape())
    bv = {(0, 0): rng.standard_normal((4, 5)),
          (2, 1): rng.standard_normal((5, 6))}
    u = D.helmholtz_solve(g, a, bf, 0.3, rhs, bv)
    import scipy.sparse as sp
    M = sp.diags(a.ravel()) - 0.3 * A
    r = (rhs + 0.3 * D._bterm(g, k, bv)).ravel()
    ref = spla.spsolve(M.tocsc(), r).reshape(g.cell_shape())
    assert np.max(np.abs(u - ref)) <= 1e-11 * np.max(np.abs(ref))


def _manufactured(n):
    """u = cos(pi x) sin(2 pi z) + cos(2 pi y) on [0,1]^3: x faces Neumann
    (du/dx = 0 at x = 0, 1), y periodic, z faces Dirichlet (u = cos(2 pi y))."""
    g = Geom.make((n, n, n), periodic=(0, 1, 0), bclo=(BC_NEUMANN, 0, BC_DIRICHLET),
                  bchi=(BC_NEUMANN, 0, BC_DIRICHLET))
    c = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    pi = np.pi
    u = np.cos(pi * X) * np.sin(2 * pi * Z) + np.cos(2 * pi * Y)
    lap = (-(pi ** 2) - 4 * pi ** 2) * np.cos(pi * X) * np.sin(2 * pi * Z) \
        - 4 * pi ** 2 * np.cos(2 * pi * Y)
    # Dirichlet values on the z faces (z = 0 and z = 1): u = cos(2 pi y)
    yy, xx = np.meshgrid(c, c, indexing="ij")
    bv = {(2, 0): np.cos(2 * pi * yy), (2, 1): np.cos(2 * pi * yy)}
    return g, u, lap, bv


def test_manufactured_solution_second_order():
    errs = []
    for n in (8, 16, 32):
        g, u, lap, bv = _manufactured(n)
        rho, beta = 2.0, 0.01
        rhs = rho * u - beta * lap
        bf = D.eta_faces(np.ones(g.cell_shape()), g.periodic)
        uh = D.helmholtz_solve(g, np.full(g.cell_shape(), rho), bf, beta, rhs, bv)
        errs.append(np.max(np.abs(uh - u)))
        # the explicit operator is a consistent Laplacian away from the
        # Dirichlet faces (first order in the boundary cells)
        L = D.apply(g, bf, u, bv)
        assert np.max(np.abs((L - lap)[1:-1])) <= 40.0 / n ** 2 * np.max(np.abs(lap))
    rates = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert min(rates) > 1.7, (errs, rates)


def test_eta_faces_definition():
    g = Geom.make((6, 5, 4), periodic=(0, 1, 1))
    c = (np.arange(6) + 0.5)
    eta = np.broadcast_to(1.0 + 0.1 * c, g.cell_shape()).copy()
    f = D.eta_faces(eta, g.periodic)
    fx = f[0][0, 0]
    # interior faces: arithmetic mean of the two cells (exact for linear data)
    assert np.allclose(fx[1:-1], 1.0 + 0.1 * np.arange(1, 6))
    # domain faces: the adjacent cell's value (fixup_eta_on_domain_faces)
    assert fx[0] == eta[0, 0, 0] and fx[-1] == eta[0, 0, -1]
    # periodic directions: face 0 and face n carry the same mean
    assert np.allclose(f[1][:, 0, :], f[1][:, -1, :])


def test_divtau_is_laplacian_over_rho():
    g = Geom.make((16, 16, 16))
    c = (np.arange(16) + 0.5) / 16
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    v = np.stack([np.sin(2 * np.pi * X), np.cos(2 * np.pi * Y) * np.sin(2 * np.pi * Z),
                  np.zeros_like(X)])
    rho = np.full(g.cell_shape(), 2.0)
    dt_ = D.divtau(g, v, rho, 0.3)
    h = 1.0 / 16
    sym = (2 * np.cos(2 * np.pi * h) - 2) / h ** 2
    assert np.allclose(dt_[0], 0.3 * sym * v[0] / 2.0, atol=1e-12)
    assert np.allclose(dt_[1], 0.3 * 2 * sym * v[1] / 2.0, atol=1e-12)
    assert np.max(np.abs(dt_[2])) == 0.0
