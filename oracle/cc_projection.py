"""Cell-centred (approximate) projection: FP64 CPU oracle.

TEST INFRASTRUCTURE ONLY (the checker for mgx_apply_cc_projection); the
product library never loads it.

incflo::ApplyCCProjection (src/projection/incflo_apply_cc_projection.cpp:199-487),
restated:
  u += (s/rho) gp                      unless incremental (:209-229)
  u -= u_old                           incremental (:231-238)
  ghost cells 0, inflow faces the inflow velocity unless incremental (:241-250)
  u_mac = MOL::ExtrapVelToFaces(u)     (:329-343)
  beta = s / face-average(rho), or s / ro_0 at constant density (:253-320)
  MAC projection of u_mac, phi from zero (:345-348), fluxes F = -beta grad phi
    (getFluxes, :362-371)
  gphi = average_mac_to_ccvel(F) = per direction the mean of the two faces (:373-380)
  u += gphi; p_cc = phi (+= incremental); gp = -gphi rho / s (-= incremental;
    rho = ro_0 at constant density) (:382-458); u += u_old when incremental.

MOL::ExtrapVelToFaces lives in AMReX-Hydro (not vendored): this build's
restatement of the published method of lines face velocity -- the normal
component extrapolated from both neighbouring cells with limited slopes and
upwinded: left state u_L = u_{i-1} + s_{i-1}/2, right state u_R = u_i - s_i/2;
the face takes u_L when the mean is positive, u_R when negative, 0 when
u_L < 0 < u_R (a rarefaction) or the mean vanishes.  Slopes: Colella's
fourth-order limited slope (the centred fourth-order difference limited by
twice the one-sided differences, with the neighbours' second-order
monotonised-central slopes) in periodic directions and away from
non-periodic faces; the monotonised-central slope one cell further in; zero
slope in the cells on a non-periodic face.  A non-periodic domain face takes
the ghost value (0 on walls, the inflow velocity on inflow faces), or on a
pressure (Dirichlet) face the adjacent cell's value.  Parity unpinned
against AMReX-Hydro's MOL extrapolation near non-periodic faces.
"""
import numpy as np

from . import BC_DIRICHLET, BC_INFLOW, mac_beta_from_density, mac_fluxes, mac_project


def _mc(dl, dr):
    dc = 0.5 * (dl + dr)
    lim = np.where(dl * dr >= 0, 2.0 * np.minimum(np.abs(dl), np.abs(dr)), 0.0)
    return np.sign(dc) * np.minimum(lim, np.abs(dc))


def slopes(q, ax, periodic):
    """limited slopes of q (cells) along axis ax"""
    n = q.shape[ax]
    sh = lambda k: np.roll(q, -k, axis=ax)
    dl, dr = q - sh(-1), sh(1) - q
    mc = _mc(dl, dr)
    dc = 0.5 * (dl + dr)
    lim = np.where(dl * dr >= 0, 2.0 * np.minimum(np.abs(dl), np.abs(dr)), 0.0)
    d4 = 4.0 / 3.0 * dc - (np.roll(mc, -1, axis=ax) + np.roll(mc, 1, axis=ax)) / 6.0
    s4 = np.sign(dc) * np.minimum(lim, np.abs(d4))
    if periodic:
        return s4
    idx = np.arange(n).reshape([-1 if a == ax else 1 for a in range(3)])
    s = np.where((idx >= 2) & (idx <= n - 3), s4, mc)
    return np.where((idx == 0) | (idx == n - 1), 0.0, s)


def mol_faces(g, ug, small=0.0):
    """u_mac (face arrays, n_d + 1 faces along d) from cell velocities with one
    ghost layer ug (3, nz+2, ny+2, nx+2).  Non-periodic domain faces: the
    ghost value (walls 0, inflow faces the inflow velocity), or the adjacent
    interior cell's value on pressure (Dirichlet) faces, where the velocity
    extrapolates outward."""
    out = []
    for d in range(3):
        ax = 2 - d
        q = ug[d, 1:-1, 1:-1, 1:-1]
        n = g.n[d]
        s = slopes(q, ax, g.periodic[d])
        uL = np.roll(q + 0.5 * s, 1, axis=ax)  # face f: left state from cell f-1
        uR = q - 0.5 * s                       # right state from cell f
        avg = 0.5 * (uL + uR)
        val = np.where(avg > small, uL, np.where(avg < -small, uR, 0.0))
        val = np.where((uL < 0) & (uR > 0), 0.0, val)
        f = np.concatenate([val, np.take(val, [0], axis=ax)], axis=ax)
        if not g.periodic[d]:
            for side in (0, 1):
                bc = g.bchi[d] if side else g.bclo[d]
                if bc == BC_DIRICHLET:
                    v = np.take(q, [n - 1 if side else 0], axis=ax)
                else:
                    gh = np.take(ug[d], [n + 1 if side else 0], axis=ax)
                    v = _interior(gh, ax)
                sl = [slice(None)] * 3
                sl[ax] = slice(n, n + 1) if side else slice(0, 1)
                f[tuple(sl)] = v
        out.append(f)
    return out


def _interior(a, ax):
    """drop the ghost layer in the two directions other than ax"""
    sl = [slice(1, -1)] * 3
    sl[ax] = slice(None)
    return a[tuple(sl)]


def cc_project(g, vel, gp, rho, scaling, vel_old=None, incremental=False, inflow=None,
               const_rho=None, rtol=1e-12):
    """Returns (u_new, gp_new_increment_or_value, p_cc, info): gp / p_cc are the
    new values (non-incremental) or the increments (incremental)."""
    nz, ny, nx = g.cell_shape()
    u = vel.copy()
    r = np.full(g.cell_shape(), const_rho) if const_rho else rho
    if not incremental:
        u = u + gp * (scaling / r)
    else:
        u = u - vel_old
    ug = np.zeros((3, nz + 2, ny + 2, nx + 2))
    ug[:, 1:-1, 1:-1, 1:-1] = u
    if inflow is not None and not incremental:
        for d in range(3):
            for side in (0, 1):
                bc = g.bchi[d] if side else g.bclo[d]
                if g.periodic[d] or bc != BC_INFLOW:
                    continue
                ax = 3 - d
                sl = [slice(None)] * 4
                sl[ax] = -1 if side else 0
                for c in range(3):
                    sl[0] = c
                    v = ug[tuple(sl)]
                    v[...] = inflow[2 * d + side][c]
    um = mol_faces(g, ug)
    if const_rho:
        beta = [np.full(um[d].shape, scaling / const_rho) for d in range(3)]
    else:
        beta = mac_beta_from_density(g, rho, scaling)
    umc = [a.copy() for a in um]
    phi, info, ret = mac_project(g, beta, umc, rtol=rtol)
    if ret != 0:
        raise RuntimeError("cc projection oracle: MAC solve did not converge")
    F = mac_fluxes(g, beta, phi)
    gphi = np.stack([0.5 * (np.take(F[d], range(0, g.n[d]), axis=2 - d) +
                            np.take(F[d], range(1, g.n[d] + 1), axis=2 - d)) for d in range(3)])
    unew = u + gphi
    if incremental:
        unew = unew + vel_old
    gpn = -gphi * r / scaling
    return unew, gpn, phi, info
