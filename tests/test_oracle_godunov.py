"""Pins the oracle's Godunov PPM advection (oracle/orc_godunov.c).

AMReX-Hydro (where the reference's Godunov lives) is not vendored, so the
restated method is pinned by what the mathematics fixes:
  * PPM parabola: exact for quadratics away from extrema; constant and
    monotone-preserving (no new extrema) on steps;
  * constant states are preserved exactly by every stage;
  * uniform translation: with constant advecting velocity the face states
    approach the exact solution at t^{n+1/2} at second order or better, and
    the convective term approaches -u.grad(s) at t^{n+1/2};
  * reflection symmetry (x -> -x, u -> -u) and the transonic (Burgers)
    rule of the normal velocity.
"""
import numpy as np
import pytest

import oracle


def _geom(n):
    return oracle.Geom.make((n, n, n))


def _cells(n):
    x = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    return X, Y, Z


def test_ppm_exact_for_quadratic_profiles(orc):
    # s = a + b x + c x^2 on a window away from the periodic seam: the
    # parabola through the cell averages is the quadratic itself, so the
    # traced value equals the exact average over the swept interval
    n = 64
    g = _geom(n)
    h = 1.0 / n
    i = np.arange(n)
    xl, xr = i * h, (i + 1) * h
    a, b, c = 0.3, 1.7, -2.1
    prim = lambda x: a * x + b * x ** 2 / 2 + c * x ** 3 / 3
    avg = (prim(xr) - prim(xl)) / h
    s = np.broadcast_to(avg, (n, n, n)).copy()
    dt = 0.4 * h
    U = 1.0
    Im, Ip = orc.ppm_predict(g, s, 0, ucell=np.full((n, n, n), U), dt=dt)
    # right-face value traced over U dt: average of s over [xr - U dt, xr]
    exact = (prim(xr) - prim(xr - U * dt)) / (U * dt)
    interior = slice(3, n - 3)
    # monotone quadratic on [0,1] only where it has no extremum: c<0 gives max at x=b/(-2c)
    xm = -b / (2 * c)
    ok = (np.abs(xl - xm) > 3 * h) & (np.abs(xr - xm) > 3 * h)
    sel = np.zeros(n, bool)
    sel[interior] = True
    sel &= ok
    assert np.max(np.abs(Ip[0, 0, sel] - exact[sel])) < 1e-12
    assert np.array_equal(Im, s)  # u > 0: the left side takes the cell value


def test_ppm_step_no_new_extrema(orc):
    n = 32
    g = _geom(n)
    s = np.zeros((n, n, n))
    s[:, :, 8:20] = 1.0
    for U in (0.7, -0.7):
        Im, Ip = orc.ppm_predict(g, s, 0, ucell=np.full((n, n, n), U), dt=0.5 / n)
        for a in (Im, Ip):
            assert a.min() >= -1e-15 and a.max() <= 1 + 1e-15


def test_constant_states_preserved(orc):
    n = 16
    g = _geom(n)
    vel = np.empty((3, n, n, n))
    vel[0], vel[1], vel[2] = 0.7, -0.3, 0.2
    dt = 0.5 / n
    um = orc.godunov_extrap_vel(g, vel, dt)
    for d, c in enumerate((0.7, -0.3, 0.2)):
        assert np.allclose(um[d], c, rtol=0, atol=1e-15)
    st = orc.godunov_edge_states(g, vel, um, dt)
    for d in range(3):
        for q, c in enumerate((0.7, -0.3, 0.2)):
            assert np.allclose(st[d, q], c, rtol=0, atol=1e-15)
    conv = orc.convective_term(g, vel, um, st)
    assert np.max(np.abs(conv)) < 1e-12


def _translation_error(orc, n):
    """Components 1 and 2 are passive profiles advected by constant (U, V, W)."""
    g = _geom(n)
    h = 1.0 / n
    X, Y, Z = _cells(n)
    U, V, W = 1.0, 0.5, -0.25
    dt = 0.5 * h
    k = 2 * np.pi
    f = lambda x, y, z: np.sin(k * x) * np.cos(k * y) + 0.5 * np.sin(k * z)
    # cell averages of a smooth profile ~ point values + O(h^2); use point values
    vel = np.empty((3, n, n, n))
    vel[0] = U
    vel[1] = V
    vel[2] = W
    um = [np.full((n, n, n), U), np.full((n, n, n), V), np.full((n, n, n), W)]
    s = f(X, Y, Z)
    q = np.stack([vel[0], s, vel[2]])
    # advected quantity = component 1 (the advecting u_mac is given)
    st = orc.godunov_edge_states(g, q, um, dt)
    th = 0.5 * dt
    errs = []
    for d, (xs, ys, zs) in enumerate(((X - h / 2, Y, Z), (X, Y - h / 2, Z), (X, Y, Z - h / 2))):
        exact = f(xs - U * th, ys - V * th, zs - W * th)
        errs.append(np.max(np.abs(st[d, 1] - exact)))
    conv = orc.convective_term(g, q, um, st)
    grad = (k * np.cos(k * (X - U * th)) * np.cos(k * (Y - V * th)) * U
            - k * np.sin(k * (X - U * th)) * np.sin(k * (Y - V * th)) * V
            + 0.5 * k * np.cos(k * (Z - W * th)) * W)
    ec = conv[1] + grad
    errs += [np.max(np.abs(ec)), np.mean(np.abs(ec))]
    return np.array(errs)


def test_translation_converges(orc):
    e1 = _translation_error(orc, 16)
    e2 = _translation_error(orc, 32)
    rate = np.log2(e1 / e2)
    # face states: second order in the max norm (the CW84 limiter clips smooth
    # extrema, so not third); the convective term differences those states:
    # first order at the clipped extrema (max norm), ~2 in the mean (L1)
    assert np.all(rate[:3] > 1.8), rate
    assert rate[3] > 0.9 and rate[4] > 1.5, rate
    assert np.all(e2[:3] < 0.02)


def test_reflection_symmetry_and_transonic_rule(orc):
    n = 16
    g = _geom(n)
    rng = np.random.default_rng(4)
    vel = rng.standard_normal((3, n, n, n))
    dt = 0.2 / n
    um = orc.godunov_extrap_vel(g, vel, dt)
    # mirror x: cell i -> n-1-i, u -> -u; x-face i -> face (n - i) mod n
    velm = vel[:, :, :, ::-1].copy()
    velm[0] *= -1
    umm = orc.godunov_extrap_vel(g, velm, dt)
    back = -np.roll(umm[0][:, :, ::-1], 1, axis=2)
    assert np.allclose(back, um[0], rtol=0, atol=1e-13)
    assert np.allclose(np.roll(umm[1][:, :, ::-1], 0, axis=2), um[1], rtol=0, atol=1e-13)
    # transonic: a diverging 1-D velocity jump gives 0 at the jump
    v2 = np.zeros((3, n, n, n))
    v2[0][:, :, : n // 2] = -1.0
    v2[0][:, :, n // 2:] = 1.0
    u2 = orc.godunov_extrap_vel(g, v2, dt)
    assert np.all(u2[0][:, :, n // 2] == 0.0)
    # converging symmetric jump: a standing shock, 0 at the face
    u3 = orc.godunov_extrap_vel(g, -v2, dt)
    assert np.all(u3[0][:, :, n // 2] == 0.0)
    # asymmetric shock (2 | -1): it moves right, the left state is upwind
    v4 = np.zeros((3, n, n, n))
    v4[0][:, :, : n // 2] = 2.0
    v4[0][:, :, n // 2:] = -1.0
    u4 = orc.godunov_extrap_vel(g, v4, dt)
    assert np.all(u4[0][:, :, n // 2] == 2.0)


# ---- scalars (density, tracers) with the MAC velocities ----------------------
def _rand_faces(rng, g, scale=1.0):
    return [scale * rng.standard_normal(g.cell_shape()) for _ in range(3)]


def test_scalar_convective_form_equals_velocity_edge_states(orc):
    # a scalar in convective form takes the same steps as a velocity component
    # in edge-state mode (the pinned path above)
    n = 16
    g = _geom(n)
    rng = np.random.default_rng(11)
    vel = rng.standard_normal((3,) + g.cell_shape())
    um = _rand_faces(rng, g, 0.5)
    dt = 0.2 / n
    ref = orc.godunov_edge_states(g, vel, um, dt)
    got = orc.godunov_scalar_states(g, vel, um, dt, iconserv=[0, 0, 0])
    assert np.array_equal(got, ref)
    conv_ref = orc.convective_term(g, vel, um, ref)
    conv = orc.scalar_advection(g, vel, um, got, iconserv=[0, 0, 0])
    assert np.array_equal(conv, conv_ref)


def test_scalar_constant_state_preserved_in_both_forms(orc):
    # u_mac not divergence free: both transverse forms vanish on a constant,
    # so the face states stay constant; conv = -s div(u) (conservative), 0 (convective)
    n = 16
    g = _geom(n)
    rng = np.random.default_rng(12)
    um = _rand_faces(rng, g, 0.7)
    s = np.full((2,) + g.cell_shape(), 2.5)
    dt = 0.3 / n
    st = orc.godunov_scalar_states(g, s, um, dt, iconserv=[1, 0])
    assert np.max(np.abs(st - 2.5)) <= 1e-14
    conv = orc.scalar_advection(g, s, um, st, iconserv=[1, 0])
    divu = sum((np.roll(um[d], -1, axis=2 - d) - um[d]) * n for d in range(3))
    assert np.max(np.abs(conv[0] + 2.5 * divu)) <= 1e-12 * np.max(np.abs(divu)) * 2.5
    assert np.max(np.abs(conv[1])) <= 1e-12 * np.max(np.abs(divu)) * 2.5


def test_scalar_conservative_form_conserves_and_reduces_to_convective(orc):
    n = 16
    g = _geom(n)
    rng = np.random.default_rng(13)
    s = 1.0 + 0.3 * rng.random((1,) + g.cell_shape())
    um = _rand_faces(rng, g, 0.6)
    dt = 0.25 / n
    st = orc.godunov_scalar_states(g, s, um, dt, iconserv=[1])
    conv = orc.scalar_advection(g, s, um, st, iconserv=[1])
    # -div(F) telescopes over the periodic domain
    assert abs(conv.sum()) <= 1e-12 * np.abs(conv).sum()
    # the conservative transverse form differs from the convective one ...
    stc = orc.godunov_scalar_states(g, s, um, dt, iconserv=[0])
    assert np.max(np.abs(st - stc)) > 1e-6
    # ... except under a uniform advecting velocity, where the two coincide
    U = [np.full(g.cell_shape(), v) for v in (0.8, -0.5, 0.3)]
    a = orc.godunov_scalar_states(g, s, U, dt, iconserv=[1])
    b = orc.godunov_scalar_states(g, s, U, dt, iconserv=[0])
    assert np.max(np.abs(a - b)) <= 1e-14


# ---- domain boundaries (inflow ext_dir, outflow foextrap, walls) ------------
EXT, FOE, HOE, REVEN, RODD, INT = 3, 2, 4, 1, -1, 0


def _ghosted(s, per, lo_fill, hi_fill):
    """s: (nc, nz, ny, nx) -> (nc, nz+4, ny+4, nx+4); periodic dims wrap,
    x (the only non-periodic dim here) takes lo_fill / hi_fill ghost columns
    (callables (c, k) -> (nc, nz+4, ny+4) values for ghost column c) or 'edge'."""
    out = np.pad(s, ((0, 0), (2, 2), (2, 2), (2, 2)), mode="wrap")
    if not per[0]:
        for c, fill in ((0, lo_fill), (1, lo_fill), (-1, hi_fill), (-2, hi_fill)):
            src = 2 if c >= 0 else -3
            out[:, :, :, c] = out[:, :, :, src] if fill == "edge" else fill(c)
    return out


def test_boundaries_constant_state_kept(orc):
    # inflow / outflow in x, periodic y, z; a constant state with matching
    # ghosts stays exactly constant on every face, both transverse forms
    n = (12, 8, 10)
    g = orc.Geom.make(n, periodic=(0, 1, 1), bclo=(3, 0, 0), bchi=(2, 0, 0))
    rng = np.random.default_rng(21)
    um = [0.5 * rng.standard_normal(orc.face_shape(g, d)) for d in range(3)]
    for d in (1, 2):  # periodic directions: the duplicate face repeats face 0
        ax = 2 - d
        sl0 = [slice(None)] * 3; sl0[ax] = 0
        sl1 = [slice(None)] * 3; sl1[ax] = -1
        um[d][tuple(sl1)] = um[d][tuple(sl0)]
    s = np.full((2,) + g.cell_shape(), 1.7)
    sg = _ghosted(s, (0, 1, 1), "edge", "edge")
    bc = np.zeros((2, 6), int)
    bc[:, 0], bc[:, 3] = EXT, FOE
    st = orc.gd_edge_states(g, sg, bc, um, 0.3 / 12, iconserv=[1, 0])
    for d in range(3):
        assert np.max(np.abs(st[d] - 1.7)) <= 1e-14, d
    conv = orc.gd_conv(g, s, um, st, iconserv=[1, 0])
    divu = sum(np.diff(um[d], axis=2 - d) * n[d] for d in range(3))
    assert np.max(np.abs(conv[0] + 1.7 * divu)) <= 1e-12 * np.abs(divu).max()
    assert np.max(np.abs(conv[1])) <= 1e-12 * np.abs(divu).max()


def _inflow_error(orc, m, U=0.8):
    """s(x, t) = f(x - U t) through an inflow face at x = 0 (ext_dir ghosts
    hold the exact boundary values at t^{n+1/2}) and out through x = 1
    (foextrap); returns the max error of the x-face states at t^{n+1/2}."""
    n = (m, 4, 4)
    g = orc.Geom.make(n, periodic=(0, 1, 1), bclo=(3, 0, 0), bchi=(2, 0, 0),
                      prob_hi=(1.0, 4.0 / m, 4.0 / m))
    h = 1.0 / m
    dt = 0.5 * h / U
    f = lambda x: np.sin(2 * np.pi * x) + 0.5 * np.cos(4 * np.pi * x)
    F = lambda x: -np.cos(2 * np.pi * x) / (2 * np.pi) + 0.5 * np.sin(4 * np.pi * x) / (4 * np.pi)
    xl = np.arange(-2, m + 2) * h
    avg = (F(xl + h) - F(xl)) / h  # cell averages at t^n, ghosts included
    s = np.broadcast_to(avg[2:-2], (1,) + g.cell_shape()).copy()
    sg = _ghosted(s, (0, 1, 1), "edge", "edge")
    sg[:, :, :, 0] = sg[:, :, :, 1] = f(0.0)  # boundary value on the inflow face
    sg[:, :, :, -2:] = avg[-2:]              # smooth continuation downstream
    um = [np.full(orc.face_shape(g, 0), U), np.zeros(orc.face_shape(g, 1)),
          np.zeros(orc.face_shape(g, 2))]
    bc = np.zeros((1, 6), int)
    bc[0, 0], bc[0, 3] = EXT, FOE
    st = orc.gd_edge_states(g, sg, bc, um, dt)
    xf = np.arange(m + 1) * h
    exact = f(xf - U * dt / 2)
    assert np.all(st[0][0, :, :, 0] == f(0.0))  # the inflow face carries the boundary value
    return np.max(np.abs(st[0][0, 0, 0, 1:] - exact[1:]))


def test_boundaries_inflow_translation_converges(orc):
    e1, e2 = _inflow_error(orc, 32), _inflow_error(orc, 64)
    assert e2 < 5e-3
    assert e1 / e2 > 3.0, (e1, e2)


def test_boundaries_walls_and_outflow_rules(orc):
    # x-lo no-slip wall (all components ext_dir 0), x-hi outflow (foextrap):
    # the wall face carries u = 0 and zero tangential states; on the outflow
    # face an inward normal velocity is clipped to 0
    n = (10, 8, 8)
    g = orc.Geom.make(n, periodic=(0, 1, 1), bclo=(3, 0, 0), bchi=(2, 0, 0))
    rng = np.random.default_rng(22)
    vel = rng.standard_normal((3,) + g.cell_shape())
    vel[0][:, :, -1] = -np.abs(vel[0][:, :, -1]) - 0.5  # flowing back into the domain at x-hi
    vg = _ghosted(vel, (0, 1, 1), lambda c: 0.0, "edge")
    bc = np.zeros((3, 6), int)
    bc[:, 0], bc[:, 3] = EXT, FOE
    dt = 0.2 / 10
    um = orc.gd_extrap_vel(g, vg, bc, dt)
    assert np.all(um[0][:, :, 0] == 0.0)
    assert np.all(um[0][:, :, -1] == 0.0)
    st = orc.gd_edge_states(g, vg, bc, um, dt)
    assert np.all(st[0][:, :, :, 0] == 0.0)
    # the outflow face takes the interior trace (not the ghost) for every component
    assert np.max(np.abs(st[0][1:, :, :, -1])) > 1e-3
