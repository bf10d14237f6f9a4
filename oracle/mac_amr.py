"""Two-level composite MAC projection (AMR, refinement ratio 2): FP64 CPU oracle.

TEST INFRASTRUCTURE ONLY (the checker of the GPU composite MAC solve); the
product library never loads it.

incflo projects the MAC velocities of all AMR levels with one MacProjector
(src/convection/incflo_compute_MAC_projected_velocities.cpp:96-133, 272-300)
and relies on it to average the fine faces down (:301-302).  AMReX's
multilevel cell-centred solver (MLABecLaplacian / MLPoisson inside MLMG, its
coarse-fine interpolation and reflux) is not vendored, so this restates the
composite problem as a conservative finite-volume scheme:

* F is the union of the fine boxes (fine cells aligned with coarse cells,
  away from non-periodic domain faces).  Unknowns: the fine cells in F and
  the coarse cells not covered by F.
* Every face between two unknowns carries the two-point flux T (phi_j - phi_i),
  T = beta * face area / distance between the two cell centres: beta_c between
  coarse cells, beta_f between fine cells, and across a fine face on the
  coarse-fine boundary beta_f with distance 1.5 h_f (fine centre to coarse
  centre along the normal).  A coarse cell next to F sees the 4 fine faces on
  each of its coarse-fine faces (reflux).  Dirichlet domain faces: distance
  h_c / 2 to the boundary value 0; Neumann faces: no flux.
* The volume-weighted operator (A phi)_i = sum T (phi_j - phi_i) is symmetric;
  the composite Laplacian is A / V.
* Right-hand side: V div u, fine cells from fine faces, uncovered coarse cells
  from coarse faces, where the coarse-fine faces take the mean of the 4 fine
  faces on them.
* Projection: solve A phi = V div u (volume-weighted mean zero when singular),
  subtract T / area * (phi difference) from every face between unknowns and
  from Dirichlet faces, then average the fine faces down onto every coarse
  face under F (its boundary included).  The composite divergence of the
  result is the solver residual.

Arrays are numpy, x fastest: coarse cells (nz, ny, nx), fine cells
(2nz, 2ny, 2nx), face arrays with n_d + 1 entries along d.  In a periodic
direction face n_d is face 0 (both entries hold the same value).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import BC_DIRICHLET


class MacAmr:
    def __init__(self, g, F):
        """g: coarse oracle.Geom; F: fine-cell mask of shape (2nz, 2ny, 2nx)."""
        self.g = g
        self.nc = (g.n[0], g.n[1], g.n[2])
        self.nf = tuple(2 * v for v in self.nc)
        self.per = [bool(g.periodic[d]) for d in range(3)]
        self.F = np.asarray(F, bool)
        Fc = self.F.reshape(self.nc[2], 2, self.nc[1], 2, self.nc[0], 2)
        self.cov = Fc.any(axis=(1, 3, 5))
        if not np.array_equal(self.cov, Fc.all(axis=(1, 3, 5))):
            raise ValueError("the fine region must consist of whole coarse cells")
        self.hc = [g.dx[d] for d in range(3)]
        self.hf = [0.5 * h for h in self.hc]
        fi = np.argwhere(self.F)
        ci = np.argwhere(~self.cov)
        self.nfu, self.ncu = len(fi), len(ci)
        self.fid = -np.ones(self.F.shape, int)
        self.fid[tuple(fi.T)] = np.arange(self.nfu)
        self.cid = -np.ones(self.cov.shape, int)
        self.cid[tuple(ci.T)] = self.nfu + np.arange(self.ncu)
        self.N = self.nfu + self.ncu
        self.vol = np.concatenate([np.full(self.nfu, np.prod(self.hf)),
                                   np.full(self.ncu, np.prod(self.hc))])
        self.singular = not any(
            not self.per[d] and (g.bclo[d] == BC_DIRICHLET or g.bchi[d] == BC_DIRICHLET)
            for d in range(3))
        self.links, self.dirichlet = self._links()

    def _area(self, h, d):
        return h[(d + 1) % 3] * h[(d + 2) % 3]

    def _face(self, d, pos, n):
        """canonical face index (d, k, j, i); periodic face n_d is face 0"""
        pos = list(pos)
        if self.per[d] and pos[d] == n[d]:
            pos[d] = 0
        return (d, pos[2], pos[1], pos[0])

    def _links(self):
        """links (a, b, geometric factor area/distance, kind, face) with a on the
        low side of the face; kind 0 coarse-coarse, 1 fine-fine, 2 coarse-fine;
        Dirichlet faces (a, factor, face, side)."""
        links, dirichlet = [], []
        nc, nf = self.nc, self.nf
        for d in range(3):
            af, ac = self._area(self.hf, d), self._area(self.hc, d)
            for (k, j, i) in np.argwhere(self.F):
                p = [i, j, k]
                a = self.fid[k, j, i]
                for s in (-1, 1):
                    q = list(p)
                    q[d] += s
                    if self.per[d]:
                        q[d] %= nf[d]
                    elif q[d] < 0 or q[d] >= nf[d]:
                        raise ValueError("fine region touches a non-periodic domain face")
                    fpos = list(p)
                    if s > 0:
                        fpos[d] += 1
                    face = self._face(d, fpos, nf)
                    b = self.fid[q[2], q[1], q[0]]
                    if b >= 0:
                        if s > 0:
                            links.append((a, b, af / self.hf[d], 1, face))
                    else:
                        c = self.cid[q[2] // 2, q[1] // 2, q[0] // 2]
                        lo, hi = (a, c) if s > 0 else (c, a)
                        links.append((lo, hi, af / (1.5 * self.hf[d]), 2, face))
            for (k, j, i) in np.argwhere(~self.cov):
                p = [i, j, k]
                a = self.cid[k, j, i]
                for s in (-1, 1):
                    q = list(p)
                    q[d] += s
                    fpos = list(p)
                    if s > 0:
                        fpos[d] += 1
                    if not self.per[d] and (q[d] < 0 or q[d] >= nc[d]):
                        bc = self.g.bclo[d] if s < 0 else self.g.bchi[d]
                        if bc == BC_DIRICHLET:
                            dirichlet.append((a, ac / (0.5 * self.hc[d]), self._face(d, fpos, nc), s))
                        continue
                    q[d] %= nc[d]
                    b = self.cid[q[2], q[1], q[0]]
                    if b >= 0 and s > 0:
                        links.append((a, b, ac / self.hc[d], 0, self._face(d, fpos, nc)))
        return links, dirichlet

    def _beta(self, kind, face, bc_, bf):
        d, k, j, i = face
        return (bc_ if kind == 0 else bf)[d][k, j, i]

    def matrix(self, bc_, bf):
        """volume-weighted composite operator A (sparse, symmetric)"""
        rows, cols, vals = [], [], []
        diag = np.zeros(self.N)
        for (a, b, gf, kind, face) in self.links:
            T = self._beta(kind, face, bc_, bf) * gf
            rows += [a, b]
            cols += [b, a]
            vals += [T, T]
            diag[a] -= T
            diag[b] -= T
        for (a, gf, face, s) in self.dirichlet:
            diag[a] -= self._beta(0, face, bc_, bf) * gf
        rows += list(range(self.N))
        cols += list(range(self.N))
        vals += list(diag)
        return sp.csr_matrix((vals, (rows, cols)), shape=(self.N, self.N))

    def _sync_periodic(self, u, n):
        for d in range(3):
            if self.per[d]:
                ax = 2 - d
                sl0 = [slice(None)] * 3
                sln = [slice(None)] * 3
                sl0[ax] = 0
                sln[ax] = n[d]
                u[d][tuple(sln)] = u[d][tuple(sl0)]
        return u

    def under_fine(self, d):
        """coarse faces of direction d with a covered cell on either side"""
        ax = 2 - d
        covp = np.pad(self.cov, [(1, 1) if a == ax else (0, 0) for a in range(3)],
                      mode="wrap" if self.per[d] else "constant")
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, -1)
        hi[ax] = slice(1, None)
        return covp[tuple(lo)] | covp[tuple(hi)]

    def average_down(self, uf, uc):
        """coarse faces under the fine level <- mean of the 4 fine faces"""
        out = []
        for d in range(3):
            ax = 2 - d
            tang = [a for a in range(3) if a != ax]
            sl = [slice(None)] * 3
            sl[ax] = slice(0, None, 2)
            fe = np.asarray(uf[d])[tuple(sl)]
            s = 0.0
            for t0 in (0, 1):
                for t1 in (0, 1):
                    idx = [slice(None)] * 3
                    idx[tang[0]] = slice(t0, None, 2)
                    idx[tang[1]] = slice(t1, None, 2)
                    s = s + fe[tuple(idx)]
            out.append(np.where(self.under_fine(d), 0.25 * s, np.asarray(uc[d], dtype=np.float64)))
        return out

    def divergence(self, uf, uc):
        """composite divergence per unknown (fine cells first)"""
        ucs = self.average_down(uf, uc)
        divf = sum(np.diff(uf[d], axis=2 - d) / self.hf[d] for d in range(3))
        divc = sum(np.diff(ucs[d], axis=2 - d) / self.hc[d] for d in range(3))
        out = np.zeros(self.N)
        out[:self.nfu] = divf[self.F]
        out[self.nfu:] = divc[~self.cov]
        return out

    def unknowns(self, phif, phic):
        return np.concatenate([phif[self.F], phic[~self.cov]])

    def levels(self, x):
        """(phi_f, phi_c) from the unknowns; covered coarse cells get the mean
        of their 8 fine cells, fine cells outside F are 0"""
        phif = np.zeros(self.F.shape)
        phif[self.F] = x[:self.nfu]
        phic = np.zeros(self.cov.shape)
        phic[~self.cov] = x[self.nfu:]
        m = sum(phif[a::2, b::2, c::2] for a in (0, 1) for b in (0, 1) for c in (0, 1)) / 8.0
        return phif, np.where(self.cov, m, phic)

    def apply(self, bc_, bf, phif, phic):
        """composite Laplacian (A / V) of (phi_f, phi_c) per unknown"""
        return self.matrix(bc_, bf) @ self.unknowns(phif, phic) / self.vol

    def project(self, bc_, bf, uc, uf):
        """returns (uc_new, uf_new, phi_f, phi_c, x, rhs) with rhs the composite
        divergence before the projection"""
        A = self.matrix(bc_, bf)
        div = self.divergence(uf, uc)
        rhs = self.vol * div
        if self.singular:
            rhs = rhs - self.vol * (rhs.sum() / self.vol.sum())
            K = sp.bmat([[A, sp.csr_matrix(self.vol[:, None])],
                         [sp.csr_matrix(self.vol[None, :]), None]]).tocsc()
            x = spla.spsolve(K, np.concatenate([rhs, [0.0]]))[:self.N]
        else:
            x = spla.spsolve(A.tocsc(), rhs)
        ufn = [np.array(a, dtype=np.float64) for a in uf]
        ucn = [np.array(a, dtype=np.float64) for a in uc]
        for (a, b, gf, kind, face) in self.links:
            d, k, j, i = face
            h = self.hc if kind == 0 else self.hf
            flux = self._beta(kind, face, bc_, bf) * gf / self._area(h, d) * (x[b] - x[a])
            (ucn if kind == 0 else ufn)[d][k, j, i] -= flux
        for (a, gf, face, s) in self.dirichlet:
            d, k, j, i = face
            g_ = self._beta(0, face, bc_, bf) * gf / self._area(self.hc, d)
            ucn[d][k, j, i] -= g_ * ((0.0 - x[a]) if s > 0 else (x[a] - 0.0))
        self._sync_periodic(ufn, self.nf)
        self._sync_periodic(ucn, self.nc)
        ucn = self.average_down(ufn, ucn)
        phif, phic = self.levels(x)
        return ucn, ufn, phif, phic, x, div


class MacAmrLevels:
    """The composite MAC problem over levels 0..L (ratio 2 between
    neighbours), MacAmr's construction applied to each pair of neighbouring
    levels: level l's unknowns are its cells in F_l not covered by F_{l+1}
    (level 0 covers the domain); a face between two unknowns of one level
    carries beta_l (phi_j - phi_i) / h_l; a level-l face on the boundary of
    F_l links its cell to the level-(l-1) cell behind it at distance 1.5 h_l
    (so a level-(l-1) cell next to F_l sees the 4 finer faces: reflux); a
    domain face of any level is Dirichlet (value 0 at h_l / 2) or Neumann.
    F_{l+1} must be properly nested in F_l (the cell behind every coarse-fine
    face is an unknown of the level below).  Right-hand side: V div u per
    unknown with the faces under a finer level replaced by the mean of its
    faces (from the finest level down); the projection subtracts each link's
    flux from its face (the finer face for coarse-fine links) and averages
    the faces down again.  Fine levels may touch Dirichlet and Neumann faces.

    g: level-0 oracle.Geom; Fs[l - 1]: level l's cell mask, shape
    (2^l nz, 2^l ny, 2^l nx)."""

    def __init__(self, g, Fs):
        self.g = g
        self.L = L = len(Fs)
        self.n = [tuple(g.n[d] << l for d in range(3)) for l in range(L + 1)]
        self.h = [[g.dx[d] / 2 ** l for d in range(3)] for l in range(L + 1)]
        self.per = [bool(g.periodic[d]) for d in range(3)]
        self.F = [np.ones((g.n[2], g.n[1], g.n[0]), bool)] + [np.asarray(F, bool) for F in Fs]
        self.cov = []
        for l in range(L + 1):
            nx, ny, nz = self.n[l]
            if l < L:
                Fc = self.F[l + 1].reshape(nz, 2, ny, 2, nx, 2)
                c = Fc.any(axis=(1, 3, 5))
                if not np.array_equal(c, Fc.all(axis=(1, 3, 5))):
                    raise ValueError("a fine region must consist of whole cells of the level below")
                if (c & ~self.F[l]).any():
                    raise ValueError("level %d reaches outside level %d" % (l + 1, l))
            else:
                c = np.zeros((nz, ny, nx), bool)
            self.cov.append(c)
        self.unk = [self.F[l] & ~self.cov[l] for l in range(L + 1)]
        self.uid, off = [], 0
        for l in range(L + 1):
            ids = -np.ones(self.unk[l].shape, int)
            cnt = int(self.unk[l].sum())
            ids[self.unk[l]] = off + np.arange(cnt)
            self.uid.append(ids)
            off += cnt
        self.N = off
        self.vol = np.concatenate([np.full(int(self.unk[l].sum()), np.prod(self.h[l]))
                                   for l in range(L + 1)])
        self.singular = not any(
            not self.per[d] and (g.bclo[d] == BC_DIRICHLET or g.bchi[d] == BC_DIRICHLET)
            for d in range(3))
        self.links, self.dirichlet = self._links()

    @staticmethod
    def _area(h, d):
        return h[(d + 1) % 3] * h[(d + 2) % 3]

    def _face(self, l, d, pos):
        """(level, d, k, j, i) of a level-l face; periodic face n_d is face 0"""
        pos = list(pos)
        if self.per[d] and pos[d] == self.n[l][d]:
            pos[d] = 0
        return (l, d, pos[2], pos[1], pos[0])

    def _links(self):
        """links (a, b, area/distance, face) with a on the low side of the
        face; Dirichlet faces (a, area/distance, face, side)"""
        links, dirichlet = [], []
        for l in range(self.L + 1):
            n, h = self.n[l], self.h[l]
            for d in range(3):
                ar = self._area(h, d)
                for (k, j, i) in np.argwhere(self.unk[l]):
                    p = [i, j, k]
                    a = self.uid[l][k, j, i]
                    for s in (-1, 1):
                        q = list(p)
                        q[d] += s
                        fpos = list(p)
                        if s > 0:
                            fpos[d] += 1
                        face = self._face(l, d, fpos)
                        if q[d] < 0 or q[d] >= n[d]:
                            if not self.per[d]:
                                bc = self.g.bclo[d] if s < 0 else self.g.bchi[d]
                                if bc == BC_DIRICHLET:
                                    dirichlet.append((a, ar / (0.5 * h[d]), face, s))
                                continue
                            q[d] %= n[d]
                        if self.F[l][q[2], q[1], q[0]]:
                            b = self.uid[l][q[2], q[1], q[0]]
                            if b >= 0 and s > 0:
                                links.append((a, b, ar / h[d], face))
                            continue  # b < 0: covered, the finer level links to a
                        c = self.uid[l - 1][q[2] // 2, q[1] // 2, q[0] // 2]
                        if c < 0:
                            raise ValueError("level %d is not properly nested in level %d" % (l, l - 1))
                        lo, hi = (a, c) if s > 0 else (c, a)
                        links.append((lo, hi, ar / (1.5 * h[d]), face))
        return links, dirichlet

    @staticmethod
    def _beta(betas, face):
        l, d, k, j, i = face
        return betas[l][d][k, j, i]

    def matrix(self, betas):
        """volume-weighted composite operator A (sparse, symmetric); betas[l][d]
        level l's face coefficients"""
        rows, cols, vals = [], [], []
        diag = np.zeros(self.N)
        for (a, b, gf, face) in self.links:
            T = self._beta(betas, face) * gf
            rows += [a, b]
            cols += [b, a]
            vals += [T, T]
            diag[a] -= T
            diag[b] -= T
        for (a, gf, face, s) in self.dirichlet:
            diag[a] -= self._beta(betas, face) * gf
        rows += list(range(self.N))
        cols += list(range(self.N))
        vals += list(diag)
        return sp.csr_matrix((vals, (rows, cols)), shape=(self.N, self.N))

    def _sync_periodic(self, u, n):
        for d in range(3):
            if self.per[d]:
                ax = 2 - d
                sl0 = [slice(None)] * 3
                sln = [slice(None)] * 3
                sl0[ax] = 0
                sln[ax] = n[d]
                u[d][tuple(sln)] = u[d][tuple(sl0)]
        return u

    def _under(self, l, d):
        ax = 2 - d
        covp = np.pad(self.cov[l], [(1, 1) if a == ax else (0, 0) for a in range(3)],
                      mode="wrap" if self.per[d] else "constant")
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, -1)
        hi[ax] = slice(1, None)
        return covp[tuple(lo)] | covp[tuple(hi)]

    def average_down(self, us):
        """faces of every level under the next finer one <- the mean of its 4
        faces on them, from the finest level down"""
        out = [[np.array(a, dtype=np.float64) for a in u] for u in us]
        for l in range(self.L - 1, -1, -1):
            for d in range(3):
                ax = 2 - d
                tang = [a for a in range(3) if a != ax]
                sl = [slice(None)] * 3
                sl[ax] = slice(0, None, 2)
                fe = out[l + 1][d][tuple(sl)]
                s = 0.0
                for t0 in (0, 1):
                    for t1 in (0, 1):
                        idx = [slice(None)] * 3
                        idx[tang[0]] = slice(t0, None, 2)
                        idx[tang[1]] = slice(t1, None, 2)
                        s = s + fe[tuple(idx)]
                out[l][d] = np.where(self._under(l, d), 0.25 * s, out[l][d])
        return out

    def divergence(self, us):
        """composite divergence per unknown (level 0 first)"""
        ua = self.average_down(us)
        parts = []
        for l in range(self.L + 1):
            dv = sum(np.diff(ua[l][d], axis=2 - d) / self.h[l][d] for d in range(3))
            parts.append(dv[self.unk[l]])
        return np.concatenate(parts)

    def unknowns(self, phis):
        return np.concatenate([phis[l][self.unk[l]] for l in range(self.L + 1)])

    def levels(self, x):
        """phi per level from the unknowns; cells under a finer level take the
        mean of its 8 cells (from the finest level down), cells outside F 0"""
        out = []
        for l in range(self.L + 1):
            p = np.zeros(self.unk[l].shape)
            p[self.unk[l]] = x[self.uid[l][self.unk[l]]]
            out.append(p)
        for l in range(self.L - 1, -1, -1):
            f = out[l + 1]
            m = sum(f[a::2, b::2, c::2] for a in (0, 1) for b in (0, 1) for c in (0, 1)) / 8.0
            out[l] = np.where(self.cov[l], m, out[l])
        return out

    def apply(self, betas, phis):
        """composite Laplacian (A / V) per unknown"""
        return self.matrix(betas) @ self.unknowns(phis) / self.vol

    def project(self, betas, us):
        """returns (us_new, phis, x, div) with div the composite divergence
        before the projection"""
        A = self.matrix(betas)
        div = self.divergence(us)
        rhs = self.vol * div
        if self.singular:
            rhs = rhs - self.vol * (rhs.sum() / self.vol.sum())
            K = sp.bmat([[A, sp.csr_matrix(self.vol[:, None])],
                         [sp.csr_matrix(self.vol[None, :]), None]]).tocsc()
            x = spla.spsolve(K, np.concatenate([rhs, [0.0]]))[:self.N]
        else:
            x = spla.spsolve(A.tocsc(), rhs)
        un = [[np.array(a, dtype=np.float64) for a in u] for u in us]
        for (a, b, gf, face) in self.links:
            l, d, k, j, i = face
            un[l][d][k, j, i] -= self._beta(betas, face) * gf / self._area(self.h[l], d) * (x[b] - x[a])
        for (a, gf, face, s) in self.dirichlet:
            l, d, k, j, i = face
            g_ = self._beta(betas, face) * gf / self._area(self.h[l], d)
            un[l][d][k, j, i] -= g_ * ((0.0 - x[a]) if s > 0 else (x[a] - 0.0))
        for l in range(self.L + 1):
            self._sync_periodic(un[l], self.n[l])
        un = self.average_down(un)
        return un, self.levels(x), x, div
