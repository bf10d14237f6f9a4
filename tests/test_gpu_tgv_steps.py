"""Ten steps of the periodic Taylor-Green vortex at 64^3 (unit cube, rho = 1,
mu = 1e-3, one max_grid_size = 32 box layout) through the whole projection
path on the GPU, against the same steps composed from the oracle's pieces.

One step, as incflo's predictor orders it (src/incflo_apply_predictor.cpp):
  forces = -gp / rho; u_mac = Godunov extrapolation of u to faces with the
  forces (incflo_compute_MAC_projected_velocities.cpp:168-178); MAC projection
  with beta = dt / rho (:272-300); convective term from the Godunov edge states
  with the projected u_mac (incflo_compute_advection_term.cpp); u* = u + dt
  (conv + mu/rho lap u - gp/rho) (the viscous term explicit here, the same on
  both sides); nodal projection with the gp add-back, sigma = dt / rho
  (incflo_apply_nodal_projection.cpp:29-267), gp = G phi.
Both sides solve to 1e-12; the velocity and gp after ten steps agree to 1e-9
relative, and the kinetic energy decays."""
import sys

import numpy as np
import pytest

from incflo_amd import boxes as B

pytestmark = pytest.mark.gpu

N = 64
STEPS = 10
MU = 1e-3


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _tgv(n):
    c = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    tp = 2 * np.pi
    u = np.sin(tp * X) * np.cos(tp * Y) * np.cos(tp * Z)
    v = -np.cos(tp * X) * np.sin(tp * Y) * np.cos(tp * Z)
    return np.stack([u, v, np.zeros_like(u)])


def _lap(u, h):
    out = np.zeros_like(u)
    for ax in (1, 2, 3):
        out += (np.roll(u, 1, axis=ax) + np.roll(u, -1, axis=ax) - 2 * u) / (h * h)
    return out


def _full_faces(um):
    """faces of n cells per direction (low faces) -> n + 1 with the periodic duplicate"""
    return [np.concatenate([um[d], np.take(um[d], [0], axis=2 - d)], axis=2 - d) for d in range(3)]


def test_taylor_green_ten_steps_match_oracle(ctx, orc):
    from incflo_amd import mgx
    n = N
    h = 1.0 / n
    dt = 0.5 * h  # CFL 0.5 at max |u| = 1
    g = orc.Geom.make((n, n, n))
    bxs = B.chop_domain((n, n, n), 32)
    lev = mgx.make_level((n, n, n), bxs)
    A = mgx.Advection(ctx, lev)
    M = mgx.MacProjector(ctx, lev)
    M.updateBeta(dt)
    P = mgx.NodalProjector(ctx, lev)
    vel = _tgv(n)
    gp = np.zeros_like(vel)
    vo, gpo = vel.copy(), gp.copy()
    beta = [np.full(orc.face_shape(g, d), dt) for d in range(3)]
    ke0 = 0.5 * np.mean(np.sum(vel * vel, axis=0))
    for step in range(STEPS):
        _say(f"case tgv 64^3 step {step}")
        # ---- GPU ----
        vd = ctx.array(B.to_boxed(vel, bxs, 3))
        fd = ctx.array(B.to_boxed(-gp, bxs, 1))
        um = [ctx.array(B.face_field_size(bxs, 1, d)) for d in range(3)]
        A.extrap_vel_to_faces(vd, 3, dt, um[0], um[1], um[2], umac_ng=1, force=fd, f_ng=1)
        im = M.project(um[0], um[1], um[2], 1, 1e-12, 0.0)
        assert im.converged
        conv = ctx.array(B.field_size(bxs, 0, 3))
        A.convective_term(vd, 3, um[0], um[1], um[2], 1, dt, conv, 0, force=fd, f_ng=1)
        c = B.from_boxed(conv.download(), bxs, 0, g.cell_shape(), ncomp=3)
        ustar = vel + dt * (c + MU * _lap(vel, h) - gp)
        vd = ctx.array(B.to_boxed(ustar, bxs, 1))
        gpd = ctx.array(B.to_boxed(gp, bxs, 0))
        pd = ctx.array(B.field_size(bxs, 0, 1, True))
        info = P.apply_nodal_projection(vd, 1, gpd, 0, pd, 0, dt, const_rho=1.0, rtol=1e-12,
                                        atol=0.0)
        assert info.converged
        vel = B.from_boxed(vd.download(), bxs, 1, g.cell_shape(), ncomp=3)
        gp = B.from_boxed(gpd.download(), bxs, 0, g.cell_shape(), ncomp=3)
        # ---- oracle ----
        force = -gpo
        umo = _full_faces(orc.godunov_extrap_vel(g, vo, dt, force=force))
        _, mi, ret = orc.mac_project(g, beta, umo, rtol=1e-12)
        assert ret == 0
        umo_n = [np.take(umo[d], range(n), axis=2 - d) for d in range(3)]
        st = orc.godunov_edge_states(g, vo, umo_n, dt, force=force)
        co = orc.convective_term(g, vo, umo_n, st)
        uso = vo + dt * (co + MU * _lap(vo, h) - gpo)
        ug = np.zeros(g.ghost_vec_shape())
        ug[:, 1:-1, 1:-1, 1:-1] = uso + dt * gpo  # gp add-back, sigma = dt / rho
        orc.fill_vel_ghosts(g, ug)
        phio, gphio, io, ret = orc.nodal_project(g, np.full(g.cell_shape(), dt), ug, rtol=1e-12)
        assert ret == 0
        vo = ug[:, 1:-1, 1:-1, 1:-1].copy()
        gpo = gphio
    sc = np.max(np.abs(vo))
    _say(f"  after {STEPS} steps: max |u - u_oracle| {np.max(np.abs(vel - vo)):.2e} of {sc:.3f}")
    assert np.max(np.abs(vel - vo)) <= 1e-9 * sc
    assert np.max(np.abs(gp - gpo)) <= 1e-9 * np.max(np.abs(gpo))
    ke = 0.5 * np.mean(np.sum(vel * vel, axis=0))
    assert ke < ke0
    A.close()
    M.close()
    P.close()


@pytest.fixture(scope="module")
def ctx():
    from incflo_amd import mgx
    c = mgx.Context(0)
    yield c
    c.close()
