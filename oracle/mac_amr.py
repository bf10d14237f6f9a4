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
