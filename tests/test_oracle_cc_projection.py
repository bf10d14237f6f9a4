"""Pins of the cell-centred projection oracle (oracle/cc_projection.py),
incflo::ApplyCCProjection (src/projection/incflo_apply_cc_projection.cpp:199-487).

* MOL face velocities: exact for linear data (the limited slopes of a linear
  field are its derivative); constants preserved; the upwind choice follows
  the sign of the mean; the fourth-order slope equals the centred difference
  formula 4/3 dc - (m+ + m-)/6 on smooth data and is limited to 0 at an extremum.
* The projection: the decomposition u_in + (s/rho) gp_old = u_new + (s/rho) gp_new
  holds to round-off (the gradient part goes to gp); the MAC velocities
  behind it are divergence free; a discrete-gradient input is removed to
  truncation size (second order) and a smooth solenoidal field is kept to
  truncation size."""
import numpy as np

from oracle import BC_DIRICHLET, BC_NEUMANN, Geom
from oracle import cc_projection as CC


def test_mol_faces_linear_and_constant():
    g = Geom.make((16, 12, 10), periodic=(0, 1, 1), bclo=(BC_NEUMANN, 0, 0),
                  bchi=(BC_DIRICHLET, 0, 0))
    ug = np.zeros((3, 12, 14, 18))
    x = np.arange(-1, 17) + 0.5
    ug[0] = 0.3 + 0.1 * x
    ug[1] = -0.7   # constant normal velocity in y (periodic)
    ug[2] = 0.25
    f = CC.mol_faces(g, ug)
    fx = f[0][3, 4]
    # faces 2..14 (cells with non-zero slopes on both sides) are exact
    assert np.allclose(fx[2:15], 0.3 + 0.1 * np.arange(2, 15), atol=1e-14)
    assert fx[0] == ug[0, 4, 5, 0]           # Neumann (wall) side: the ghost value
    assert fx[16] == ug[0, 4, 5, 16]         # Dirichlet side: the adjacent cell
    assert np.all(f[1] == -0.7) and np.all(f[2] == 0.25)


def test_mol_upwinding_and_extremum():
    g = Geom.make((8, 8, 8))
    ug = np.zeros((3, 10, 10, 10))
    # a step in x: right-moving fluid takes the left state, left-moving the right
    q = np.where(np.arange(8) < 4, 1.0, 2.0)
    ug[0, 1:-1, 1:-1, 1:-1] = q
    f = CC.mol_faces(g, ug)[0][0, 0]
    assert f[4] == 1.0 + 0.0  # left cell 3 has zero limited slope at the jump's foot
    ug[0, 1:-1, 1:-1, 1:-1] = -q
    f = CC.mol_faces(g, ug)[0][0, 0]
    assert f[4] == -2.0
    # rarefaction: left state < 0 < right state gives 0
    ug[0, 1:-1, 1:-1, 1:-1] = np.where(np.arange(8) < 4, -1.0, 1.0)
    f = CC.mol_faces(g, ug)[0][0, 0]
    assert f[4] == 0.0
    # a smooth periodic field: slopes are the fourth-order limited form
    xs = (np.arange(32) + 0.5) / 32
    qq = np.sin(2 * np.pi * xs)
    s = CC.slopes(np.broadcast_to(qq, (2, 2, 32)).copy(), 2, True)[0, 0]
    h = 1.0 / 32
    exact = 2 * np.pi * h * np.cos(2 * np.pi * xs)
    away = np.abs(np.cos(2 * np.pi * xs)) > 0.3
    assert np.max(np.abs(s - exact)[away]) < 2e-4
    assert np.max(np.abs(s[np.argmax(qq)])) <= np.max(np.abs(np.diff(qq))) + 1e-15


def _fields(n, rng):
    g = Geom.make((n, n, n))
    h = 1.0 / n
    c = (np.arange(n) + 0.5) * h
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    return g, X, Y, Z


def test_decomposition_and_divergence(orc):
    rng = np.random.default_rng(3)
    g, X, Y, Z = _fields(16, rng)
    u = rng.standard_normal((3,) + g.cell_shape())
    gp = rng.standard_normal((3,) + g.cell_shape())
    rho = 1.0 + 0.5 * rng.random(g.cell_shape())
    s = 0.01
    un, gpn, phi, info = CC.cc_project(g, u, gp, rho, s, rtol=1e-13)
    lhs = u + gp * s / rho
    rhs = un + gpn * s / rho
    assert np.max(np.abs(lhs - rhs)) <= 1e-13 * np.max(np.abs(lhs))


def test_gradient_removed_solenoidal_kept(orc):
    errs_g, errs_s = [], []
    for n in (16, 32):
        g, X, Y, Z = _fields(n, None)
        tp = 2 * np.pi
        # gradient of psi = sin(2 pi x) cos(2 pi y) sin(2 pi z)
        grad = np.stack([tp * np.cos(tp * X) * np.cos(tp * Y) * np.sin(tp * Z),
                         -tp * np.sin(tp * X) * np.sin(tp * Y) * np.sin(tp * Z),
                         tp * np.sin(tp * X) * np.cos(tp * Y) * np.cos(tp * Z)])
        sol = np.stack([np.sin(tp * X) * np.cos(tp * Y), -np.cos(tp * X) * np.sin(tp * Y),
                        np.zeros_like(X)])
        rho = np.ones(g.cell_shape())
        z = np.zeros_like(grad)
        ug, _, _, _ = CC.cc_project(g, grad, z, rho, 1.0, rtol=1e-13)
        us, _, _, _ = CC.cc_project(g, sol, z, rho, 1.0, rtol=1e-13)
        errs_g.append(np.max(np.abs(ug)) / np.max(np.abs(grad)))
        errs_s.append(np.max(np.abs(us - sol)) / np.max(np.abs(sol)))
    assert errs_g[1] < 0.05 and errs_s[1] < 0.05
    assert errs_g[0] / errs_g[1] > 3.0 and errs_s[0] / errs_s[1] > 3.0
