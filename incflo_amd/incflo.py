"""Device-resident incflo time stepping on one periodic level.

The host side here is control flow only: every field lives in device memory
in the boxed layout of include/mgx.h and every arithmetic step is a C-ABI
call into libmgx.so.  It mirrors incflo's driver for the Godunov path with
constant density, no tracers and implicit viscosity solved per component:

  InitData (src/incflo.cpp:40-100): InitialProjection (src/setup/init.cpp:450-493)
    and InitialIterations (:402-447);
  Evolve / Advance (src/incflo.cpp:138-197, src/incflo_advance.cpp): one
    ApplyPredictor per step (src/incflo_apply_predictor.cpp) with the fixed
    dt of incflo.fixed_dt (src/incflo_compute_dt.cpp:265-280).

ApplyPredictor, in the reference's order:
  copy new -> old velocity, FillPatch of the old velocity (periodic images);
  divtau_o = div(mu grad u_old) / rho                      mgx_diff_apply
  vel_forces = -gp / rho + divtau_o, ghosts filled          mgx_compute_vel_forces
  u_mac = ExtrapVelToFaces(u_old, vel_forces)               mgx_adv_extrap_vel_to_faces
  MAC projection, beta = dt / face-average(rho)             mgx_mac_project
  conv = Godunov convective term                            mgx_adv_convective_term
  forces again with rho_nph (no divtau), u* = u_old + dt (conv + forces)
                                                            mgx_update_velocity
  (rho - dt div(mu grad)) u = rho u*                        mgx_diff_solve
  nodal projection with scaling dt (incremental in the initial iterations)
                                                            mgx_apply_nodal_projection
"""
import numpy as np

from . import boxes as B
from . import mgx

NG_STATE = 3   # nghost_state() for Godunov (src/incflo.H:714-724)
NG_FORCE = 1   # nghost_force() for Godunov
NG_MAC = 1     # nghost_mac() for Godunov without EB (src/incflo.H:736-741)


class DeviceIncflo:
    """Periodic single level, constant density ro_0, viscosity mu."""

    def __init__(self, ctx, n, prob_hi, mu, velocity, ro_0=1.0, max_grid_size=32,
                 mac_rtol=1e-11, mac_atol=1e-14, nodal_rtol=1e-11, nodal_atol=1e-14,
                 diff_rtol=1e-11, diff_atol=1e-14, mode=mgx.MODE_FP64):
        self.ctx = ctx
        self.n = tuple(n)
        self.bxs = B.chop_domain(self.n, max_grid_size)
        self.lev = mgx.make_level(self.n, self.bxs, prob_hi=prob_hi)
        self.mu = mu
        self.tol = dict(mac=(mac_rtol, mac_atol), nodal=(nodal_rtol, nodal_atol),
                        diff=(diff_rtol, diff_atol))
        self.mode = mode
        bx = self.bxs
        self.vel = ctx.array(B.to_boxed(velocity, bx, NG_STATE))
        self.vel_o = ctx.array(self.vel.n)
        self.rho = ctx.array(B.to_boxed(np.full((n[2], n[1], n[0]), float(ro_0)), bx, 0))
        self.gp = ctx.array(B.field_size(bx, 0, 3))
        self.p = ctx.array(B.field_size(bx, 0, 1, True))
        self.conv = ctx.array(B.field_size(bx, 0, 3))
        self.divtau = ctx.array(B.field_size(bx, 0, 3))
        self.forces = ctx.array(B.field_size(bx, NG_FORCE, 3))
        self.umac = [ctx.array(B.face_field_size(bx, NG_MAC, d)) for d in range(3)]
        self.step = mgx.Step(ctx, self.lev)
        self.adv = mgx.Advection(ctx, self.lev)
        self.mac = mgx.MacProjector(ctx, self.lev)
        self.nodal = mgx.NodalProjector(ctx, self.lev)
        self.diff = mgx.Diffusion(ctx, self.lev)
        self.diff.set_eta(const_eta=mu)
        self.time = 0.0
        self.nstep = 0
        self.dt_last = -1.0
        self.solves = dict(mac=0, nodal=0, diff=0)
        self.step.fill_boundary(self.vel, NG_STATE, 3)

    # ---- incflo::ApplyNodalProjection -------------------------------------
    def _nodal(self, scaling, incremental):
        r, a = self.tol["nodal"]
        self.nodal.apply_nodal_projection(
            self.vel, NG_STATE, self.gp, 0, self.p, 0, scaling, density=self.rho, rho_ng=0,
            vel_old=self.vel_o if incremental else None, vel_old_ng=NG_STATE,
            incremental=incremental, rtol=r, atol=a, mode=self.mode)
        self.solves["nodal"] += 1
        self.step.fill_boundary(self.vel, NG_STATE, 3)

    def initial_projection(self):
        self._nodal(1.0, False)
        zero = mgx.lib().mgx_memset_d
        self.ctx.check(zero(self.ctx.h, self.p.ptr, 0, self.p.nbytes))
        self.ctx.check(zero(self.ctx.h, self.gp.ptr, 0, self.gp.nbytes))

    # ---- incflo::ApplyPredictor -------------------------------------------
    def predictor(self, dt, incremental=False):
        S = self.step
        self.vel_o.copy_from(self.vel)  # copy_from_new_to_old_velocity (ghosts included)
        self.diff.apply(self.vel_o, NG_STATE, 3, self.divtau, 0, rho=self.rho)
        S.compute_vel_forces(self.gp, 0, self.rho, 0, self.forces, NG_FORCE, divtau=self.divtau)
        S.fill_boundary(self.forces, NG_FORCE, 3)
        self.adv.extrap_vel_to_faces(self.vel_o, NG_STATE, dt, *self.umac, umac_ng=NG_MAC,
                                     force=self.forces, f_ng=NG_FORCE)
        self.mac.set_beta_from_density(self.rho, 0, dt)
        r, a = self.tol["mac"]
        self.mac.project(*self.umac, NG_MAC, r, a, mode=self.mode)
        self.solves["mac"] += 1
        self.adv.convective_term(self.vel_o, NG_STATE, *self.umac, NG_MAC, dt, self.conv, 0,
                                 force=self.forces, f_ng=NG_FORCE)
        # update_velocity: forces with rho_nph (= rho: constant density), no divtau
        S.compute_vel_forces(self.gp, 0, self.rho, 0, self.forces, NG_FORCE)
        S.update_velocity(self.vel, NG_STATE, self.vel_o, NG_STATE, self.conv, 0, self.forces,
                          NG_FORCE, dt)
        r, a = self.tol["diff"]
        self.diff.solve(self.vel, NG_STATE, 3, dt, acoef=self.rho, rtol=r, atol=a, mode=self.mode)
        self.solves["diff"] += 1
        self._nodal(dt, incremental)

    def initial_iterations(self, count, dt):
        for _ in range(count):
            self.predictor(dt, incremental=True)
            self.vel.copy_from(self.vel_o)  # copy_from_old_to_new_velocity

    def advance(self, dt):
        self.predictor(dt, incremental=False)
        self.time += dt
        self.nstep += 1
        self.dt_last = dt

    def run(self, stop_time, dt, initial_iterations):
        self.initial_projection()
        self.initial_iterations(initial_iterations, dt)
        while not (self.time >= stop_time - 1e-12 * dt):
            self.advance(dt)
        self.ctx.sync()

    # ---- output and restart (src/utilities/io.cpp) ---------------------------
    def write_checkpoint(self, path):
        """incflo::WriteCheckPointFile (:70-111): Header, then velocity,
        density, gradp and the pressure of the projection in use as VisMF
        multifabs under Level_0.  The reference writes p_cc when it runs the
        nodal projection and p_nd with the cell-centred one (:99-105), so a
        restart loses the pressure the run used; this build writes the
        nodal-projection pressure as p_nd."""
        mgx.write_checkpoint_header(path, [self.lev], self.nstep, self.time, self.dt_last,
                                    self.dt_last, self.dt_last)
        pre = path + "/Level_0/"
        mgx.write_multifab(self.ctx, pre + "velocity", self.lev, self.vel, NG_STATE, 3)
        mgx.write_multifab(self.ctx, pre + "density", self.lev, self.rho, 0, 1)
        mgx.write_multifab(self.ctx, pre + "gradp", self.lev, self.gp, 0, 3)
        mgx.write_multifab(self.ctx, pre + "p_nd", self.lev, self.p, 0, 1, nodal=True)

    def restart(self, path):
        """incflo::ReadCheckpointFile (:113-239) on this level's box list."""
        h = mgx.read_checkpoint_header(path)
        mine = [(tuple(lo), tuple(hi)) for lo, hi in self.bxs]
        if h["finest_level"] != 0 or h["boxes"][0] != mine:
            raise mgx.MgxError(mgx.MGX_ERR_ARG, "restart: the checkpoint's boxes differ")
        pre = path + "/Level_0/"
        mgx.read_multifab(self.ctx, pre + "velocity", self.lev, self.vel, NG_STATE, 3)
        mgx.read_multifab(self.ctx, pre + "density", self.lev, self.rho, 0, 1)
        mgx.read_multifab(self.ctx, pre + "gradp", self.lev, self.gp, 0, 3)
        mgx.read_multifab(self.ctx, pre + "p_nd", self.lev, self.p, 0, 1, nodal=True)
        self.nstep, self.time, self.dt_last = h["nstep"], h["cur_time"], h["dt"]

    def write_plotfile(self, path):
        """incflo::WritePlotFile's default variables velx, vely, velz, gpx,
        gpy, gpz, density (src/utilities/io.cpp:283-460), as one 7-component
        cell multifab."""
        n = self.n
        shape = (n[2], n[1], n[0])
        v = self.velocity()
        gp = B.from_boxed(self.gp.download(), self.bxs, 0, shape, ncomp=3)
        rho = B.from_boxed(self.rho.download(), self.bxs, 0, shape)
        data = np.concatenate([v, gp, rho[None]])
        d = self.ctx.array(B.to_boxed(data, self.bxs, 0))
        mgx.write_plotfile(self.ctx, path, [self.lev], [d], 0,
                           ["velx", "vely", "velz", "gpx", "gpy", "gpz", "density"], self.time,
                           level_steps=[self.nstep])
        d.free()

    # ---- results (host copies, for checking) -------------------------------
    def velocity(self):
        return B.from_boxed(self.vel.download(), self.bxs, NG_STATE,
                            (self.n[2], self.n[1], self.n[0]), ncomp=3)

    def pressure(self):
        return B.from_boxed(self.p.download(), self.bxs, 0, (self.n[2], self.n[1], self.n[0]),
                            nodal=True)

    def close(self):
        for h in (self.step, self.adv, self.mac, self.nodal, self.diff):
            h.close()
