"""Pins of the Robin MAC oracle (oracle/mac_robin.py): its limits are the
Dirichlet and Neumann faces of the C oracle (orc_mac.c), a manufactured
solution with general a, b, f converges at second order, and the projected
velocity is divergence free."""
import numpy as np
import pytest

from oracle import mac_robin as R

BCP, BCD, BCN, BCR = 0, 1, 2, 4


def _robin_planes(g, rng, sides, kind):
    rob = {}
    for (d, s) in sides:
        shp = R.plane_shape(g, d)
        if kind == "dirichlet":
            a, b, f = np.ones(shp), np.zeros(shp), np.zeros(shp)
        elif kind == "neumann":
            a, b, f = np.zeros(shp), np.ones(shp), np.zeros(shp)
        else:
            a = rng.random(shp)
            b = 0.1 + rng.random(shp)
            f = rng.standard_normal(shp)
        rob[(d, s)] = (a, b, f)
    return rob


@pytest.mark.parametrize("kind,bc", [("dirichlet", BCD), ("neumann", BCN)])
def test_robin_limits_equal_dirichlet_and_neumann(orc, kind, bc):
    n = (8, 6, 5)
    rng = np.random.default_rng(1)
    per = (0, 0, 1)
    gr = orc.Geom.make(n, periodic=per, bclo=(BCR, BCN, 0), bchi=(BCR, BCR, 0),
                       prob_hi=(1.0, 0.7, 0.5))
    gc = orc.Geom.make(n, periodic=per, bclo=(bc, BCN, 0), bchi=(bc, bc, 0),
                       prob_hi=(1.0, 0.7, 0.5))
    beta = [0.5 + rng.random(orc.face_shape(gr, d)) for d in range(3)]
    beta[2][-1] = beta[2][0]  # periodic z: face n is face 0
    rob = _robin_planes(gr, rng, [(0, 0), (0, 1), (1, 1)], kind)
    phi = rng.standard_normal(gr.cell_shape())
    got = R.robin_apply(gr, beta, rob, phi)
    ref = orc.mac_apply(gc, beta, phi)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    F = R.robin_fluxes(gr, beta, rob, phi)
    Fr = orc.mac_fluxes(gc, beta, phi)  # getFluxes sign: -beta grad phi
    for d in range(3):
        assert np.max(np.abs(F[d] + Fr[d])) <= 1e-12 * np.max(np.abs(Fr[d])), d


def test_robin_operator_symmetric_negative_definite(orc):
    n = (6, 5, 4)
    rng = np.random.default_rng(2)
    g = orc.Geom.make(n, periodic=(0, 0, 0), bclo=(BCR, BCN, BCR), bchi=(BCR, BCR, BCN))
    beta = [0.5 + rng.random(orc.face_shape(g, d)) for d in range(3)]
    rob = _robin_planes(g, rng, [(0, 0), (0, 1), (1, 1), (2, 0)], "general")
    A, _ = R._operator(g, beta, rob)
    A = A.toarray()
    assert np.max(np.abs(A - A.T)) <= 1e-12 * np.max(np.abs(A))
    assert np.max(np.linalg.eigvalsh(A)) < 0


def _manufactured(orc, m):
    """phi = cos(pi x) cos(2 pi y) (1 + z^2) on [0,1]^3, beta = 1, Robin on
    all six faces with a, b varying along the face; returns the max error."""
    n = (m, m, m)
    g = orc.Geom.make(n, periodic=(0, 0, 0), bclo=(BCR,) * 3, bchi=(BCR,) * 3)
    h = 1.0 / m
    xc = (np.arange(m) + 0.5) * h
    xf = np.arange(m + 1) * h
    Z, Y, X = np.meshgrid(xc, xc, xc, indexing="ij")

    def phi(x, y, z):
        return np.cos(np.pi * x) * np.cos(2 * np.pi * y) * (1 + z * z)

    def grad(x, y, z):
        return (-np.pi * np.sin(np.pi * x) * np.cos(2 * np.pi * y) * (1 + z * z),
                -2 * np.pi * np.cos(np.pi * x) * np.sin(2 * np.pi * y) * (1 + z * z),
                np.cos(np.pi * x) * np.cos(2 * np.pi * y) * 2 * z)

    lap = (-(np.pi ** 2) - 4 * np.pi ** 2) * phi(X, Y, Z) + \
        np.cos(np.pi * X) * np.cos(2 * np.pi * Y) * 2
    beta = [np.ones(orc.face_shape(g, d)) for d in range(3)]
    rob = {}
    for d in range(3):
        # face-plane coordinates (numpy order of the two remaining axes)
        P, Q = np.meshgrid(xc, xc, indexing="ij")
        for s in (0, 1):
            pos = 0.0 if s == 0 else 1.0
            if d == 0:
                x, y, z = np.full_like(P, pos), Q, P
            elif d == 1:
                x, y, z = Q, np.full_like(P, pos), P
            else:
                x, y, z = Q, P, np.full_like(P, pos)
            a = 1.0 + 0.5 * np.sin(3 * P + Q)
            b = 0.2 + 0.1 * np.cos(P - 2 * Q)
            dn = grad(x, y, z)[d] * (1.0 if s else -1.0)
            f = a * phi(x, y, z) + b * dn
            rob[(d, s)] = (a, b, f)
    x = R.robin_solve(g, beta, rob, lap)
    return np.max(np.abs(x - phi(X, Y, Z)))


def test_robin_manufactured_second_order(orc):
    e1, e2 = _manufactured(orc, 8), _manufactured(orc, 16)
    assert e2 < 0.05
    assert e1 / e2 > 3.2, (e1, e2)


def test_robin_projection_divergence_free(orc):
    n = (8, 8, 6)
    rng = np.random.default_rng(3)
    g = orc.Geom.make(n, periodic=(0, 0, 1), bclo=(BCN, BCN, 0), bchi=(BCR, BCN, 0))
    beta = [0.5 + rng.random(orc.face_shape(g, d)) for d in range(3)]
    beta[2][-1] = beta[2][0]
    shp = R.plane_shape(g, 0)
    a = np.zeros(shp)
    a[:, : shp[1] // 2] = 1.0  # outflow on the lower half, inflow above
    rob = {(0, 1): (a, 1.0 - a, np.zeros(shp))}
    u = [rng.standard_normal(orc.face_shape(g, d)) for d in range(3)]
    u[2][-1] = u[2][0]
    un, phi = R.robin_project(g, beta, rob, u)
    div = orc.mac_divergence(g, un)
    assert np.max(np.abs(div)) <= 1e-10 * np.max(np.abs(orc.mac_divergence(g, u)))
    # Neumann part of the mixed face keeps its velocity; the outflow part changes
    top = a == 0
    assert np.array_equal(un[0][:, :, -1][top], u[0][:, :, -1][top])
    assert np.max(np.abs(un[0][:, :, -1][~top] - u[0][:, :, -1][~top])) > 1e-3
