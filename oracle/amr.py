"""Two-level composite nodal projection (AMR, refinement ratio 2): FP64 CPU oracle.

TEST INFRASTRUCTURE ONLY (the checker for the GPU composite solve); the
product library never loads it.

The reference projects all AMR levels at once
(src/projection/incflo_apply_nodal_projection.cpp:124-215: one NodalProjector
over levels 0..finest_level) and averages gp down afterwards (:258-266).
AMReX's coarse-fine nodal synchronisation inside MLNodeLaplacian is not
vendored, so this restates the composite problem as the conforming Q1 finite
element discretisation on the composite mesh, which is what that
synchronisation approximates:

* F is the union of the fine boxes (cells of the fine index space, aligned
  with coarse cells).  Unknowns are the fine nodes inside F (every in-domain
  cell around them is fine) and the coarse nodes elsewhere, including the
  coarse nodes on the coarse-fine boundary dF.  Fine nodes on dF are not
  unknowns: coincident ones take the coarse value, the others the trilinear
  interpolation of the coarse face / edge they lie on.
* Elements are the fine cells of F and the coarse cells outside F.  A coarse
  Q1 element equals the Galerkin product of its 8 fine children, so the
  composite operator is P^T K P of the uniformly fine operator with the coarse
  coefficient copied to the children (tests/test_oracle_amr.py checks this).
* Rows are kept in finite-difference units of their own level:
    fine unknown m:   (A x)_m = (L_f x~_f)_m
    coarse unknown n: (A x)_n = (L_c^out x_c)_n + R(y)_n,  y = L_f x~_f on dF
  with x~_f the fine field on the closure of F, L_c^out the coarse operator
  with the coefficient of covered cells set to 0, and R the full-weighting
  restriction (P^T / 8).  The right-hand side is formed the same way from the
  nodal divergence of the fine velocity inside F and the coarse velocity
  outside F.
* After the solve: fine cells u -= sigma G_f phi~, coarse cells outside F
  u -= sigma G_c phi; covered coarse cells receive the average of their 8
  children for u and G phi (average_down, :258-266); coarse nodes under F
  take the fine value (injection).  For singular problems the mean of the
  coarse phi (after that injection) is removed.
"""
import numpy as np

from . import (Geom, fill_vel_ghosts, nodal_apply, nodal_divergence, nodal_gradient,
               nodal_prolong_add, nodal_restrict)


def refine(gc):
    g = Geom()
    for d in range(3):
        g.n[d] = 2 * gc.n[d]
        g.periodic[d] = gc.periodic[d]
        g.bclo[d] = gc.bclo[d]
        g.bchi[d] = gc.bchi[d]
        g.dx[d] = 0.5 * gc.dx[d]
    return g


def _node_any_all(g, cells):
    """For a boolean cell array: per node, whether any / all of its in-domain
    adjacent cells are set."""
    a = cells.copy()
    b = cells.copy()
    for ax, d in ((0, 2), (1, 1), (2, 0)):
        if g.periodic[d]:
            ap = np.concatenate([np.take(a, [-1], axis=ax), a], axis=ax)
            bp = np.concatenate([np.take(b, [-1], axis=ax), b], axis=ax)
            n = a.shape[ax]
            lo = [slice(None)] * 3
            hi = [slice(None)] * 3
            lo[ax] = slice(0, n)
            hi[ax] = slice(1, n + 1)
        else:
            pad = [(0, 0)] * 3
            pad[ax] = (1, 1)
            ap = np.pad(a, pad, constant_values=False)
            bp = np.pad(b, pad, constant_values=True)
            n = a.shape[ax] + 1
            lo = [slice(None)] * 3
            hi = [slice(None)] * 3
            lo[ax] = slice(0, n)
            hi[ax] = slice(1, n + 1)
        a = ap[tuple(lo)] | ap[tuple(hi)]
        b = bp[tuple(lo)] & bp[tuple(hi)]
    return a, b


class AmrGeom:
    """Coarse geometry gc, fine geometry gf = refine(gc), fine boxes
    (fine cell index lo[3], hi[3], inclusive; lo even, hi odd)."""

    OUT, DOF, BDRY = 0, 1, 2

    def __init__(self, gc, fine_boxes):
        self.gc = gc
        self.gf = refine(gc)
        self.boxes = [tuple(int(v) for v in b) for b in fine_boxes]
        fm = np.zeros(self.gf.cell_shape(), bool)
        for b in self.boxes:
            lo, hi = b[:3], b[3:]
            for d in range(3):
                assert lo[d] % 2 == 0 and hi[d] % 2 == 1, "fine boxes must align with coarse cells"
                assert 0 <= lo[d] <= hi[d] < self.gf.n[d]
            fm[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
        self.fmask = fm                      # fine cells in F
        nz, ny, nx = gc.cell_shape()
        self.cov = fm.reshape(nz, 2, ny, 2, nx, 2).any(axis=(1, 3, 5))  # coarse cells under F
        anyf, allf = _node_any_all(self.gf, fm)
        t = np.zeros(self.gf.node_shape(), np.uint8)
        t[anyf & allf] = self.DOF
        t[anyf & ~allf] = self.BDRY
        self.ftype = t
        _, allc = _node_any_all(gc, self.cov)
        self.cdof = ~allc                    # coarse unknowns

    # --- pieces -----------------------------------------------------------
    def fill_fine(self, xc, xf):
        """fine field on the closure of F: unknowns from xf, dF from xc"""
        pxc = nodal_prolong_add(self.gf, xc, np.zeros(self.gf.node_shape()))
        return np.where(self.ftype == self.DOF, xf, np.where(self.ftype == self.BDRY, pxc, 0.0))

    def inject(self, xc, xf):
        """coarse field everywhere: unknowns from xc, coarse nodes under F
        from the coincident fine node"""
        out = xc.copy()
        fine_at = xf[tuple(slice(0, None, 2) for _ in range(3))]
        fine_at = fine_at[:out.shape[0], :out.shape[1], :out.shape[2]]
        return np.where(self.cdof, out, fine_at)

    def apply(self, sig_c, sig_f, xc, xf):
        """composite operator (finite-difference units of each level)"""
        xt = self.fill_fine(xc, xf)
        yf = nodal_apply(self.gf, np.where(self.fmask, sig_f, 0.0), xt)
        yc = nodal_apply(self.gc, np.where(self.cov, 0.0, sig_c), np.where(self.cdof, xc, 0.0))
        rc = nodal_restrict(self.gf, np.where(self.ftype == self.BDRY, yf, 0.0))
        return (np.where(self.cdof, yc + rc, 0.0), np.where(self.ftype == self.DOF, yf, 0.0))

    def divergence(self, uc, uf, inflow_vel=None):
        """composite nodal divergence of the coarse (outside F) and fine
        (inside F) cell velocities (3 comps, no ghosts)"""
        ugc = np.zeros(self.gc.ghost_vec_shape())
        ugc[:, 1:-1, 1:-1, 1:-1] = np.where(self.cov[None], 0.0, uc)
        fill_vel_ghosts(self.gc, ugc, inflow_vel)
        dc = nodal_divergence(self.gc, ugc)
        ugf = np.zeros(self.gf.ghost_vec_shape())
        ugf[:, 1:-1, 1:-1, 1:-1] = np.where(self.fmask[None], uf, 0.0)
        fill_vel_ghosts(self.gf, ugf)
        df = nodal_divergence(self.gf, ugf)
        rc = nodal_restrict(self.gf, np.where(self.ftype == self.BDRY, df, 0.0))
        return (np.where(self.cdof, dc + rc, 0.0), np.where(self.ftype == self.DOF, df, 0.0))

    def average_down(self, vf):
        """mean of the 8 fine children of every coarse cell (component-wise)"""
        nz, ny, nx = self.gc.cell_shape()
        lead = vf.shape[:-3]
        return vf.reshape(lead + (nz, 2, ny, 2, nx, 2)).mean(axis=(-5, -3, -1))

    def singular(self):
        g = self.gc
        return all(g.periodic[d] or (g.bclo[d] != 1 and g.bchi[d] != 1) for d in range(3))

    # --- solve ------------------------------------------------------------
    def solve(self, sig_c, sig_f, bc, bf, rtol=1e-12, maxiter=20000):
        """Conjugate gradients on the symmetric form (coarse rows x 8);
        returns (xc, xf, iterations, 0 on success)"""
        dof_c = self.cdof.copy()
        dof_f = self.ftype == self.DOF
        # Dirichlet coarse nodes are fixed at 0
        g = self.gc
        nz, ny, nx = g.node_shape()
        dmask = np.zeros(g.node_shape(), bool)
        for d, ax in ((0, 2), (1, 1), (2, 0)):
            if g.periodic[d]:
                continue
            sl = [slice(None)] * 3
            if g.bclo[d] == 1:
                sl[ax] = 0
                dmask[tuple(sl)] = True
            if g.bchi[d] == 1:
                sl[ax] = -1
                dmask[tuple(sl)] = True
        dof_c &= ~dmask

        def A(xc, xf):
            yc, yf = self.apply(sig_c, sig_f, xc, xf)
            return -8.0 * np.where(dof_c, yc, 0.0), -np.where(dof_f, yf, 0.0)

        rc = -8.0 * np.where(dof_c, bc, 0.0)
        rf = -np.where(dof_f, bf, 0.0)
        sing = self.singular()
        ndof = dof_c.sum() + dof_f.sum()
        if sing:
            mu = (rc.sum() + rf.sum()) / ndof
            rc = np.where(dof_c, rc - mu, 0.0)
            rf = np.where(dof_f, rf - mu, 0.0)
        bn = max(np.abs(rc).max(), np.abs(rf).max())
        xc = np.zeros_like(rc)
        xf = np.zeros_like(rf)
        if bn == 0.0:
            return xc, xf, 0, 0
        # Jacobi preconditioner: the diagonal of the symmetric form
        dc, df = self._diag(sig_c, sig_f, dof_c, dof_f)
        zc, zf = np.where(dof_c, rc / dc, 0.0), np.where(dof_f, rf / df, 0.0)
        pc, pf = zc.copy(), zf.copy()
        rz = np.sum(rc * zc) + np.sum(rf * zf)
        for it in range(1, maxiter + 1):
            qc, qf = A(pc, pf)
            alpha = rz / (np.sum(pc * qc) + np.sum(pf * qf))
            xc += alpha * pc
            xf += alpha * pf
            rc -= alpha * qc
            rf -= alpha * qf
            if max(np.abs(rc).max(), np.abs(rf).max()) <= rtol * bn:
                return xc, xf, it, 0
            zc, zf = np.where(dof_c, rc / dc, 0.0), np.where(dof_f, rf / df, 0.0)
            rz2 = np.sum(rc * zc) + np.sum(rf * zf)
            pc = zc + (rz2 / rz) * pc
            pf = zf + (rz2 / rz) * pf
            rz = rz2
        return xc, xf, maxiter, 1

    def _diag(self, sig_c, sig_f, dof_c, dof_f):
        """diagonal of the symmetric form by probing with a coloured basis:
        unknowns s apart in every direction do not interact (couplings reach
        one node of their own level), s chosen so periodic wrap keeps them apart"""
        dc = np.zeros(self.gc.node_shape())
        df = np.zeros(self.gf.node_shape())
        s = next((q for q in (3, 4, 5, 6, 7, 8)
                  if all((not self.gc.periodic[d]) or self.gc.n[d] % q == 0 for d in range(3))), None)
        if s is None:
            return np.ones_like(dc), np.ones_like(df)
        for lev in (0, 1):
            shape = dc.shape if lev == 0 else df.shape
            for oz in range(s):
                for oy in range(s):
                    for ox in range(s):
                        e = np.zeros(shape)
                        e[oz::s, oy::s, ox::s] = 1.0
                        if lev == 0:
                            ec = np.where(dof_c, e, 0.0)
                            yc, yf = self.apply(sig_c, sig_f, ec, np.zeros_like(df))
                            dc += np.where(ec > 0, -8.0 * yc, 0.0)
                        else:
                            ef = np.where(dof_f, e, 0.0)
                            yc, yf = self.apply(sig_c, sig_f, np.zeros_like(dc), ef)
                            df += np.where(ef > 0, -yf, 0.0)
        dc = np.where(dof_c & (dc > 0), dc, 1.0)
        df = np.where(dof_f & (df > 0), df, 1.0)
        return dc, df

    def project(self, sig_c, sig_f, uc, uf, rtol=1e-12, inflow_vel=None):
        """composite projection; returns (uc_new, uf_new, phi_c, phi_f, gp_c, gp_f, its)
        with phi_c on every coarse node, phi_f on the closure of F (0 elsewhere),
        gp_f on fine cells of F (0 elsewhere)"""
        bc, bf = self.divergence(uc, uf, inflow_vel)
        xc, xf, its, ret = self.solve(sig_c, sig_f, bc, bf, rtol=rtol)
        assert ret == 0, "oracle CG did not converge"
        phf = self.fill_fine(xc, xf)
        phc = self.inject(xc, phf)
        if self.singular():
            m = phc.mean()
            phc = phc - m
            phf = np.where(self.ftype != self.OUT, phf - m, 0.0)
        gf = nodal_gradient(self.gf, phf) * self.fmask[None]
        gc = nodal_gradient(self.gc, phc)
        gc = np.where(self.cov[None], self.average_down(gf), gc)
        ufn = np.where(self.fmask[None], uf - sig_f[None] * gf, 0.0)
        ucn = uc - sig_c[None] * gc
        ucn = np.where(self.cov[None], self.average_down(ufn), ucn)
        return ucn, ufn, phc, phf, gc, gf, its


# ---- synchronisation (ratio 2) ------------------------------------------------
def average_down_faces(fine_cells_mask, ff, fc):
    """average_down_faces (src/convection/incflo_compute_advection_term.cpp:559-569)
    on canonical face arrays: fc[d] (coarse, n_c + 1 along d) gets, on every
    coarse face whose 4 fine faces belong to the fine level, their mean.
    fine_cells_mask: (nz, ny, nx) bool of the fine level's cells; ff[d]: fine
    faces (n_f + 1 along d).  Returns new coarse arrays."""
    out = [np.array(a, dtype=np.float64) for a in fc]
    F = np.asarray(fine_cells_mask, bool)
    for d in range(3):
        ax = 2 - d
        # a fine face belongs to the fine level when a fine cell lies on either side
        pad = [(0, 0)] * 3
        pad[ax] = (1, 1)
        Fp = np.pad(F, pad)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, -1)
        hi[ax] = slice(1, None)
        has = Fp[tuple(lo)] | Fp[tuple(hi)]
        # the 4 fine faces of each coarse face: even index along d, pairs across
        sl = [slice(None)] * 3
        sl[ax] = slice(0, None, 2)
        f_ev = ff[d][tuple(sl)]
        h_ev = has[tuple(sl)]
        tang = [a for a in range(3) if a != ax]
        s = 0.0
        ok = True
        for t0 in (0, 1):
            for t1 in (0, 1):
                idx = [slice(None)] * 3
                idx[tang[0]] = slice(t0, None, 2)
                idx[tang[1]] = slice(t1, None, 2)
                s = s + f_ev[tuple(idx)]
                ok = ok & h_ev[tuple(idx)]
        out[d] = np.where(ok, 0.25 * s, out[d])
    return out


def average_down_cells(fine_cells_mask, uf, uc):
    """average_down of a cell field: coarse cell = mean of its 8 fine cells
    where those are fine cells; uf (nz_f, ny_f, nx_f), uc (nz_c, ny_c, nx_c)."""
    F = np.asarray(fine_cells_mask, bool)
    s = sum(uf[a::2, b::2, c::2] for a in (0, 1) for b in (0, 1) for c in (0, 1))
    cov = F[0::2, 0::2, 0::2]
    return np.where(cov, 0.125 * s, uc)
