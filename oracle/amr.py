"""Composite nodal projection over AMR levels (refinement ratio 2): FP64 CPU oracle.
AmrGeom is the two-level form; AmrLevels (below) applies the same construction
to any number of levels.

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


# ---- any number of levels (ratio 2 between neighbours) ---------------------------
def dirichlet_nodes(g):
    """boolean node array: nodes on a Dirichlet domain face (fixed at 0)"""
    m = np.zeros(g.node_shape(), bool)
    for d, ax in ((0, 2), (1, 1), (2, 0)):
        if g.periodic[d]:
            continue
        sl = [slice(None)] * 3
        if g.bclo[d] == 1:
            sl[ax] = 0
            m[tuple(sl)] = True
        if g.bchi[d] == 1:
            sl[ax] = -1
            m[tuple(sl)] = True
    return m


class AmrLevels:
    """Composite nodal problem over levels 0..L, each refining the one below
    by 2 (incflo projects levels 0..finest_level in one NodalProjector,
    src/projection/incflo_apply_nodal_projection.cpp:124-215).

    The two-level construction of AmrGeom applied level by level: F_l is the
    union of level l's boxes (level 0 covers the domain), C_l the level-l cells
    under F_{l+1}.  Level-l unknowns are the nodes of F_l's interior (every
    in-domain cell around them in F_l) that are not surrounded by C_l, minus
    Dirichlet nodes.  Level-l nodes on dF_l take the trilinear interpolation
    of the level l-1 field; level-l nodes surrounded by C_l are represented on
    level l+1.  Elements: the cells of F_l \\ C_l on each level.  Rows, in
    finite-difference units of their own level:
        (A x)_l = L_l^out x~_l + R(L_{l+1}^out x~_{l+1} on dF_{l+1})
    with L_l^out the level-l operator over the cells of F_l \\ C_l.  F_{l+1}
    must be properly nested in F_l: every level-l node touching C_l lies in
    F_l's interior.  Fine boxes may touch Dirichlet and Neumann domain faces
    (ghost velocity 0 there) but not inflow faces."""

    OUT, DOF, BDRY = 0, 1, 2

    def __init__(self, gc, level_boxes):
        self.g = [gc]
        for _ in level_boxes:
            self.g.append(refine(self.g[-1]))
        self.L = L = len(level_boxes)
        self.boxes = [[tuple(int(v) for v in b) for b in bl] for bl in level_boxes]
        self.fmask = [np.ones(gc.cell_shape(), bool)]
        for l, bl in enumerate(self.boxes, 1):
            g = self.g[l]
            fm = np.zeros(g.cell_shape(), bool)
            for b in bl:
                lo, hi = b[:3], b[3:]
                for d in range(3):
                    assert lo[d] % 2 == 0 and hi[d] % 2 == 1, "boxes must align with coarser cells"
                    assert 0 <= lo[d] <= hi[d] < g.n[d]
                    if not g.periodic[d]:
                        assert not (lo[d] == 0 and g.bclo[d] == 3) and \
                            not (hi[d] == g.n[d] - 1 and g.bchi[d] == 3), \
                            "fine boxes must stay off inflow faces"
                fm[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
            self.fmask.append(fm)
        self.cov = []
        for l in range(L + 1):
            if l < L:
                nz, ny, nx = self.g[l].cell_shape()
                c = self.fmask[l + 1].reshape(nz, 2, ny, 2, nx, 2).any(axis=(1, 3, 5))
                assert not (c & ~self.fmask[l]).any(), "level %d outside level %d" % (l + 1, l)
            else:
                c = np.zeros(self.g[l].cell_shape(), bool)
            self.cov.append(c)
        self.ftype = [None]
        for l in range(1, L + 1):
            anyf, allf = _node_any_all(self.g[l], self.fmask[l])
            t = np.zeros(self.g[l].node_shape(), np.uint8)
            t[anyf & allf] = self.DOF
            t[anyf & ~allf] = self.BDRY
            self.ftype.append(t)
        self.allcov, self.dof, self.dmask = [], [], []
        for l in range(L + 1):
            anyc, allc = _node_any_all(self.g[l], self.cov[l])
            if 1 <= l < L:
                assert not (anyc & (self.ftype[l] != self.DOF)).any(), \
                    "level %d is not properly nested in level %d" % (l + 1, l)
            base = np.ones(self.g[l].node_shape(), bool) if l == 0 else self.ftype[l] == self.DOF
            self.allcov.append(allc & base)
            self.dmask.append(dirichlet_nodes(self.g[l]))
            self.dof.append(base & ~allc & ~self.dmask[l])
        self.keep = [self.fmask[l] & ~self.cov[l] for l in range(L + 1)]

    def singular(self):
        g = self.g[0]
        return all(g.periodic[d] or (g.bclo[d] != 1 and g.bchi[d] != 1) for d in range(3))

    def composite(self, xs):
        """per level: the unknowns of xs, dF_l interpolated from level l-1, 0 elsewhere"""
        out = [np.where(self.dof[0], xs[0], 0.0)]
        for l in range(1, self.L + 1):
            p = nodal_prolong_add(self.g[l], out[l - 1], np.zeros(self.g[l].node_shape()))
            t = self.ftype[l]
            out.append(np.where(self.dof[l], xs[l],
                                np.where(t == self.BDRY, np.where(self.dmask[l], 0.0, p), 0.0)))
        return out

    def _combine(self, y):
        out = []
        for l in range(self.L + 1):
            r = y[l]
            if l < self.L:
                r = r + nodal_restrict(self.g[l + 1], np.where(self.ftype[l + 1] == self.BDRY,
                                                               y[l + 1], 0.0))
            out.append(np.where(self.dof[l], r, 0.0))
        return out

    def apply(self, sigs, xs):
        """composite operator (finite-difference units of each level)"""
        xt = self.composite(xs)
        return self._combine([nodal_apply(self.g[l], np.where(self.keep[l], sigs[l], 0.0), xt[l])
                              for l in range(self.L + 1)])

    def divergence(self, us, inflow_vel=None):
        """composite nodal divergence of the cell velocities of every level
        (3 comps, no ghosts; level l's cells outside F_l \\ C_l are ignored)"""
        d = []
        for l in range(self.L + 1):
            g = self.g[l]
            ug = np.zeros(g.ghost_vec_shape())
            ug[:, 1:-1, 1:-1, 1:-1] = np.where(self.keep[l][None], us[l], 0.0)
            fill_vel_ghosts(g, ug, inflow_vel if l == 0 else None)
            d.append(nodal_divergence(g, ug))
        return self._combine(d)

    def average_down(self, vf):
        nz, ny, nx = vf.shape[-3] // 2, vf.shape[-2] // 2, vf.shape[-1] // 2
        lead = vf.shape[:-3]
        return vf.reshape(lead + (nz, 2, ny, 2, nx, 2)).mean(axis=(-5, -3, -1))

    def solve(self, sigs, bs, rtol=1e-12, maxiter=20000):
        """Jacobi-preconditioned CG on the symmetric form (level-l rows x 8^(L-l));
        returns (xs, iterations, 0 on success)"""
        L = self.L
        sc = [8.0 ** (L - l) for l in range(L + 1)]
        dof = self.dof

        def A(xs):
            ys = self.apply(sigs, xs)
            return [-sc[l] * np.where(dof[l], ys[l], 0.0) for l in range(L + 1)]

        r = [-sc[l] * np.where(dof[l], bs[l], 0.0) for l in range(L + 1)]
        ndof = sum(int(m.sum()) for m in dof)
        if self.singular():
            mu = sum(v.sum() for v in r) / ndof
            r = [np.where(dof[l], r[l] - mu, 0.0) for l in range(L + 1)]
        bn = max(np.abs(v).max() for v in r)
        xs = [np.zeros_like(v) for v in r]
        if bn == 0.0:
            return xs, 0, 0
        dg = self._diag(sigs, sc)
        z = [np.where(dof[l], r[l] / dg[l], 0.0) for l in range(L + 1)]
        p = [v.copy() for v in z]
        rz = sum(np.sum(a * b) for a, b in zip(r, z))
        for it in range(1, maxiter + 1):
            q = A(p)
            alpha = rz / sum(np.sum(a * b) for a, b in zip(p, q))
            for l in range(L + 1):
                xs[l] += alpha * p[l]
                r[l] -= alpha * q[l]
            if max(np.abs(v).max() for v in r) <= rtol * bn:
                return xs, it, 0
            z = [np.where(dof[l], r[l] / dg[l], 0.0) for l in range(L + 1)]
            rz2 = sum(np.sum(a * b) for a, b in zip(r, z))
            p = [z[l] + (rz2 / rz) * p[l] for l in range(L + 1)]
            rz = rz2
        return xs, maxiter, 1

    def _diag(self, sigs, sc):
        """diagonal of the symmetric form by probing every level with nodes s
        apart (their couplings reach one node of their own level)"""
        L = self.L
        dg = [np.zeros(self.g[l].node_shape()) for l in range(L + 1)]
        s = next((q for q in (3, 4, 5, 6, 7, 8)
                  if all((not self.g[0].periodic[d]) or self.g[0].n[d] % q == 0 for d in range(3))),
                 None)
        if s is None:
            return [np.ones_like(v) for v in dg]
        zero = [np.zeros_like(v) for v in dg]
        for l in range(L + 1):
            for oz in range(s):
                for oy in range(s):
                    for ox in range(s):
                        e = np.zeros(dg[l].shape)
                        e[oz::s, oy::s, ox::s] = 1.0
                        e = np.where(self.dof[l], e, 0.0)
                        xs = list(zero)
                        xs[l] = e
                        y = self.apply(sigs, xs)[l]
                        dg[l] += np.where(e > 0, -sc[l] * y, 0.0)
        return [np.where(self.dof[l] & (dg[l] > 0), dg[l], 1.0) for l in range(L + 1)]

    def project(self, sigs, us, rtol=1e-12, inflow_vel=None):
        """composite projection; returns (us_new, phis, gps, its): phi on every
        node of level 0 and on F_l's closure for l >= 1 (levels' nodes under
        the next finer level hold its value), G phi and the new velocity on
        every level's cells in F_l (cells under the next level: the average of
        their 8 children, applied from the finest level down)"""
        L = self.L
        bs = self.divergence(us, inflow_vel)
        xs, its, ret = self.solve(sigs, bs, rtol=rtol)
        assert ret == 0, "oracle CG did not converge"
        ph = self.composite(xs)
        for l in range(L - 1, -1, -1):
            f = ph[l + 1][::2, ::2, ::2][:ph[l].shape[0], :ph[l].shape[1], :ph[l].shape[2]]
            ph[l] = np.where(self.allcov[l], f, ph[l])
        if self.singular():
            m = ph[0].mean()
            ph[0] = ph[0] - m
            for l in range(1, L + 1):
                ph[l] = np.where(self.ftype[l] != self.OUT, ph[l] - m, 0.0)
        gp, un = [], []
        for l in range(L + 1):
            gl = nodal_gradient(self.g[l], ph[l]) * self.fmask[l][None]
            gp.append(gl)
            un.append(np.where(self.fmask[l][None], us[l] - sigs[l][None] * gl, 0.0))
        for l in range(L - 1, -1, -1):
            c = self.cov[l][None]
            gp[l] = np.where(c, self.average_down(gp[l + 1]), gp[l])
            un[l] = np.where(c, self.average_down(un[l + 1]), un[l])
        return un, ph, gp, its
