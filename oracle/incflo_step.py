"""incflo's time step on a periodic single level, composed from the oracle's
pieces: FP64 CPU oracle of the whole predictor path.

TEST INFRASTRUCTURE ONLY (the checker for the device step and for the
reference's own convergence numbers); the product library never loads it.

The sequence restated from the reference (Godunov advection, constant
density, no tracers, implicit viscosity with separate component solves,
use_mac_phi_in_godunov = false):

  InitData (src/incflo.cpp:40-100): InitialProjection, then
  InitialIterations.
  InitialProjection (src/setup/init.cpp:450-493): nodal projection of the
    initial velocity with scaling factor 1 (sigma = 1/rho), not incremental;
    p and gp are then reset to zero.
  InitialIterations (:402-447): `initial_iterations` times ApplyPredictor
    with an incremental projection, after each of which velocity (and
    density) are reset to the old values: only p and gp evolve.
  ApplyPredictor (src/incflo_apply_predictor.cpp):
    divtau_o = div(mu grad u_old) / rho            (need_divtau(): Godunov
                                                   includes diffusion in its
                                                   forcing, src/incflo.H:755)
    vel_forces = -gp / rho + divtau_o              (compute_vel_forces,
      src/incflo_compute_forces.cpp:135-148; the divtau term,
      src/convection/incflo_compute_MAC_projected_velocities.cpp:24-57)
    u_mac = ExtrapVelToFaces(u_old, vel_forces), MAC projection with
      beta = dt / face-average(rho) (:70-300)
    conv = Godunov convective term with u_mac and vel_forces
      (src/convection/incflo_compute_advection_term.cpp)
    u* = u_old + dt (conv - gp / rho)              (update_velocity,
      Implicit, src/incflo_update_velocity.cpp:67-76)
    (rho - dt div(mu grad)) u = rho u*             (diffuse_velocity, :322-334)
    nodal projection (src/projection/incflo_apply_nodal_projection.cpp:29-93)
      with scaling factor dt: non-incremental u += dt gp / rho first and
      gp = G phi, p = phi after; incremental u -= u_old first, u += u_old
      after, gp += G phi, p += phi.

Viscosity: the reference default is the tensor solve (use_tensor_solve,
src/incflo.H:392), whose operator div(mu (grad u + grad u^T)) differs from
the component solves by mu grad(div u), which the approximate projection
keeps at truncation-error size; this restatement uses the component solves
(use_tensor_solve = false), as the device step does.

Exact solution of the decaying Taylor vortex (probtype 2) and the error norms
the reference prints: src/derive/incflo_error.cpp:58-109 and
src/utilities/io.cpp:517-592 (Norm0 = max |e|, Norm2 = sqrt(sum e^2 / N); the
pressure is the cell average of p_nd minus its mean, against the exact
pressure at t - dt/2).
"""
import numpy as np

from . import (Geom, convective_term, fill_vel_ghosts, godunov_edge_states, godunov_extrap_vel,
               mac_project, nodal_project)
from . import diffusion as D

PI = 3.1415926535897932


def taylor_vortex_velocity(n, dx, t=0.0, mu=0.0):
    """probtype 2 (src/prob/prob_init_fluid.cpp:290-313 at t = 0; the exact
    solution of src/derive/incflo_error.cpp:75-101 at time t)."""
    x = (np.arange(n[0]) + 0.5) * dx[0]
    y = (np.arange(n[1]) + 0.5) * dx[1]
    Y, X = np.meshgrid(y, x, indexing="ij")
    om = PI * PI * mu
    dec = np.exp(-2.0 * om * t)
    u2 = 1.0 - np.cos(PI * (X - t)) * np.sin(PI * (Y - t)) * dec
    v2 = 1.0 + np.sin(PI * (X - t)) * np.cos(PI * (Y - t)) * dec
    nz = n[2]
    u = np.broadcast_to(u2, (nz,) + u2.shape).copy()
    v = np.broadcast_to(v2, (nz,) + v2.shape).copy()
    return np.stack([u, v, np.zeros_like(u)])


def taylor_vortex_pressure(n, dx, t_p, mu):
    """exact p at t_p (= time - dt/2 in the reference's error routine)."""
    x = (np.arange(n[0]) + 0.5) * dx[0]
    y = (np.arange(n[1]) + 0.5) * dx[1]
    Y, X = np.meshgrid(y, x, indexing="ij")
    om = PI * PI * mu
    p2 = -0.25 * (np.cos(2 * PI * (X - t_p)) + np.cos(2 * PI * (Y - t_p))) * np.exp(-4 * om * t_p)
    return np.broadcast_to(p2, (n[2],) + p2.shape).copy()


def node_to_cell(p):
    """amrex::average_node_to_cellcenter on a periodic nodal field (nz, ny, nx)."""
    s = np.zeros_like(p)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                s += np.roll(p, (-dz, -dy, -dx), axis=(0, 1, 2))
    return s / 8.0


def norms(e):
    return float(np.max(np.abs(e))), float(np.sqrt(np.sum(e * e) / e.size))


def errors(n, dx, vel, p_nd, time, dt, mu):
    """(u Norm0, u Norm2, p Norm0, p Norm2) as io.cpp prints them."""
    ue = vel[0] - taylor_vortex_velocity(n, dx, time, mu)[0]
    pc = node_to_cell(p_nd)
    pc = pc - pc.mean()
    pe = pc - taylor_vortex_pressure(n, dx, time - 0.5 * dt, mu)
    return norms(ue) + norms(pe)


def _full_faces(um):
    """n low faces per direction -> n + 1 with the periodic duplicate"""
    return [np.concatenate([um[d], np.take(um[d], [0], axis=2 - d)], axis=2 - d)
            for d in range(3)]


def _low_faces(um, n):
    return [np.take(um[d], range(n[d]), axis=2 - d) for d in range(3)]


class Incflo:
    """Periodic single-level incflo state (constant density ro_0)."""

    def __init__(self, n, prob_hi, mu, ro_0=1.0, rtol=1e-11, atol=1e-14):
        self.n = tuple(n)
        self.g = Geom.make(n, prob_hi=prob_hi)
        self.dx = [self.g.dx[d] for d in range(3)]
        self.mu = mu
        self.rho = np.full(self.g.cell_shape(), float(ro_0))
        self.rtol, self.atol = rtol, atol
        self.vel = taylor_vortex_velocity(self.n, self.dx)
        self.gp = np.zeros_like(self.vel)
        self.p = np.zeros(self.g.node_shape())
        self.time = 0.0
        self.nstep = 0

    # ---- nodal projection (incflo::ApplyNodalProjection) -----------------
    def _project(self, vel_old, scaling, incremental):
        g = self.g
        u = self.vel.copy()
        if not incremental:
            u += self.gp * (scaling / self.rho)
        else:
            u -= vel_old
        ug = np.zeros(g.ghost_vec_shape())
        ug[:, 1:-1, 1:-1, 1:-1] = u
        fill_vel_ghosts(g, ug)
        phi, gphi, info, ret = nodal_project(g, scaling / self.rho, ug, rtol=self.rtol,
                                             atol=self.atol)
        if ret != 0:
            raise RuntimeError("oracle nodal projection did not converge")
        u = ug[:, 1:-1, 1:-1, 1:-1].copy()
        if incremental:
            u += vel_old
            self.gp = self.gp + gphi
            self.p = self.p + phi
        else:
            self.gp = gphi
            self.p = phi
        self.vel = u

    def initial_projection(self):
        self._project(None, 1.0, False)
        self.p = np.zeros_like(self.p)
        self.gp = np.zeros_like(self.gp)

    # ---- ApplyPredictor ----------------------------------------------------
    def predictor(self, dt, incremental=False):
        g = self.g
        vo = self.vel.copy()
        rho = self.rho
        divtau = D.divtau(g, vo, rho, self.mu)
        forces = -self.gp / rho + divtau
        um = _full_faces(godunov_extrap_vel(g, vo, dt, force=forces))
        rf = D.eta_faces(rho, g.periodic)  # face average of rho
        beta = [dt / rf[d] for d in range(3)]
        _, mi, ret = mac_project(g, beta, um, rtol=self.rtol, atol=self.atol)
        if ret != 0:
            raise RuntimeError("oracle MAC projection did not converge")
        umn = _low_faces(um, self.n)
        st = godunov_edge_states(g, vo, umn, dt, force=forces)
        conv = convective_term(g, vo, umn, st)
        ustar = vo + dt * (conv - self.gp / rho)
        self.vel = D.diffuse_velocity(g, ustar, rho, self.mu, dt)
        self._project(vo, dt, incremental)
        return vo

    def initial_iterations(self, count, dt):
        for _ in range(count):
            vo = self.predictor(dt, incremental=True)
            self.vel = vo

    def advance(self, dt):
        self.predictor(dt, incremental=False)
        self.time += dt
        self.nstep += 1

    def run(self, stop_time, dt, initial_iterations):
        """InitData + Evolve with a fixed dt (incflo.fixed_dt: every step,
        the last one included, takes dt; src/incflo_compute_dt.cpp:265-280),
        stopping once time >= stop_time - 1e-12 dt (src/incflo.cpp:190-193)."""
        self.initial_projection()
        self.initial_iterations(initial_iterations, dt)
        while not (self.time >= stop_time - 1e-12 * dt):
            self.advance(dt)
        return errors(self.n, self.dx, self.vel, self.p, self.time, dt, self.mu)
