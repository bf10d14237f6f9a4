"""Pins the composite nodal oracle over any number of levels (oracle.amr.AmrLevels).

The construction is fixed by the conforming Q1 finite-element space on the
composite mesh, so these checks are what that definition fixes:
  * two levels: the same operator, divergence and projection as the pinned
    two-level oracle (AmrGeom);
  * three levels, periodic and bounded (fine levels on a Dirichlet and a
    Neumann face): the composite operator and divergence equal the Galerkin
    forms P^T L_u P and P^T D_u u~ of the uniformly finest operator, with an
    interpolation matrix P assembled here level by level, independently of
    the oracle's fill and restriction code;
  * the symmetric form is symmetric and annihilates constants (periodic);
  * the projection solves the composite system, applies each level's
    gradient, averages down from the finest level and injects phi.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from oracle.amr import AmrGeom, AmrLevels

# periodic 8^3 base; level 1 one box, level 2 an L shape nested one level-1 cell inside
PER = dict(n=(8, 8, 8), per=(1, 1, 1), bclo=(0, 0, 0), bchi=(0, 0, 0),
           levels=[[(2, 2, 2, 11, 11, 9)],
                   [(6, 8, 6, 13, 15, 13), (14, 6, 8, 19, 11, 15)]])
# x bounded (Dirichlet low, Neumann high), both fine levels touching both faces
BND = dict(n=(8, 8, 8), per=(0, 1, 1), bclo=(1, 0, 0), bchi=(2, 0, 0),
           levels=[[(0, 2, 2, 9, 11, 9), (12, 4, 4, 15, 9, 9)],
                   [(0, 6, 6, 13, 15, 13), (26, 10, 10, 31, 17, 17)]])


def _levels(case):
    gc = oracle.Geom.make(case["n"], periodic=case["per"], bclo=case["bclo"], bchi=case["bchi"],
                          prob_hi=(1.0, 1.0, 1.0))
    return AmrLevels(gc, case["levels"])


def _interp(A):
    """P: all levels' node arrays (concatenated; only unknowns used) -> every
    node of the finest level, built level by level from the unknowns and the
    trilinear weights of the node's position in the coarser level"""
    offs = np.cumsum([0] + [int(np.prod(g.node_shape())) for g in A.g])
    ntot = int(offs[-1])
    Q = None
    for l, g in enumerate(A.g):
        nz, ny, nx = g.node_shape()
        rows, cols, vals = [], [], []
        for k in range(nz):
            for j in range(ny):
                for i in range(nx):
                    r = i + nx * (j + ny * k)
                    if A.dof[l][k, j, i]:
                        rows.append(r); cols.append(offs[l] + r); vals.append(1.0)
                        continue
                    if l == 0 or A.dmask[l][k, j, i] or A.allcov[l][k, j, i]:
                        continue  # fixed at 0, or represented on the next level
                    gc = A.g[l - 1]
                    cz, cy, cx = gc.node_shape()
                    parts = []
                    for f, nn in ((i, cx), (j, cy), (k, cz)):
                        lo = f // 2
                        parts.append([(lo % nn, 1.0)] if f % 2 == 0 else
                                     [(lo % nn, 0.5), ((lo + 1) % nn, 0.5)])
                    for (I, wx) in parts[0]:
                        for (J, wy) in parts[1]:
                            for (K, wz) in parts[2]:
                                cr = I + cx * (J + cy * K)
                                assert not A.allcov[l - 1][K, J, I], "interpolation from a covered node"
                                row = Q.getrow(cr)
                                for c, v in zip(row.indices, row.data):
                                    rows.append(r); cols.append(c); vals.append(wx * wy * wz * v)
        Q = sp.csr_matrix((vals, (rows, cols)), shape=(nz * ny * nx, ntot))
    return Q, offs


def _finest_sigma(A, sigs):
    L = A.L
    out = np.zeros(A.g[L].cell_shape())
    for l in range(L + 1):
        f = 2 ** (L - l)
        up = np.repeat(np.repeat(np.repeat(np.where(A.keep[l], sigs[l], 0.0), f, 0), f, 1), f, 2)
        kp = np.repeat(np.repeat(np.repeat(A.keep[l], f, 0), f, 1), f, 2)
        out = np.where(kp, up, out)
    return out


def _operator_matrix(g, sig):
    """the finest-level operator as a sparse matrix, probed with nodes 4 apart"""
    shape = g.node_shape()
    n = int(np.prod(shape))
    rows, cols, vals = [], [], []
    idx = np.arange(n).reshape(shape)
    for oz in range(4):
        for oy in range(4):
            for ox in range(4):
                e = np.zeros(shape)
                e[oz::4, oy::4, ox::4] = 1.0
                y = oracle.nodal_apply(g, sig, e)
                src = np.zeros(shape, dtype=np.int64) - 1
                src[oz::4, oy::4, ox::4] = idx[oz::4, oy::4, ox::4]
                # every node's probed neighbour (at most one within distance 1)
                own = np.full(shape, -1, dtype=np.int64)
                for dz in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            s = src
                            for ax, dd in ((0, dz), (1, dy), (2, dx)):
                                per = g.periodic[2 - ax]
                                s = np.roll(s, dd, axis=ax)
                                if not per and dd != 0:
                                    sl = [slice(None)] * 3
                                    sl[ax] = 0 if dd == 1 else -1
                                    s[tuple(sl)] = -1
                            own = np.where(s >= 0, s, own)
                nz = (y != 0) & (own >= 0)
                rows.extend(idx[nz].tolist()); cols.extend(own[nz].tolist()); vals.extend(y[nz].tolist())
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def _rand_sigs(A, seed):
    rng = np.random.default_rng(seed)
    return [0.5 + rng.random(g.cell_shape()) for g in A.g]


def _split(A, v, offs):
    return [v[offs[l]:offs[l + 1]].reshape(A.g[l].node_shape()) for l in range(A.L + 1)]


@pytest.mark.parametrize("case", [PER, BND], ids=["periodic", "bounded"])
def test_three_level_operator_is_galerkin(orc, case):
    A = _levels(case)
    for l in (1, 2):
        assert A.dof[l].sum() > 0 and (A.ftype[l] == A.BDRY).sum() > 0
    assert A.allcov[0].sum() > 0 and A.allcov[1].sum() > 0
    sigs = _rand_sigs(A, 1)
    P, offs = _interp(A)
    Lu = _operator_matrix(A.g[2], _finest_sigma(A, sigs))
    rng = np.random.default_rng(2)
    xs = [np.where(A.dof[l], rng.standard_normal(A.g[l].node_shape()), 0.0) for l in range(3)]
    x = np.concatenate([v.ravel() for v in xs])
    full = _split(A, P.T @ (Lu @ (P @ x)), offs)
    ref = [np.where(A.dof[l], full[l] / 8.0 ** (2 - l), 0.0) for l in range(3)]
    ys = A.apply(sigs, xs)
    scale = max(np.abs(r).max() for r in ref)
    for l in range(3):
        assert np.abs(ys[l] - ref[l]).max() < 1e-12 * scale, l


@pytest.mark.parametrize("case", [PER, BND], ids=["periodic", "bounded"])
def test_three_level_divergence_is_galerkin(orc, case):
    A = _levels(case)
    rng = np.random.default_rng(3)
    us = [rng.standard_normal((3,) + g.cell_shape()) for g in A.g]
    L = A.L
    uu = np.zeros((3,) + A.g[L].cell_shape())
    for l in range(L + 1):
        f = 2 ** (L - l)
        up = np.repeat(np.repeat(np.repeat(us[l], f, 1), f, 2), f, 3)
        kp = np.repeat(np.repeat(np.repeat(A.keep[l], f, 0), f, 1), f, 2)
        uu = np.where(kp[None], up, uu)
    ug = np.zeros(A.g[L].ghost_vec_shape())
    ug[:, 1:-1, 1:-1, 1:-1] = uu
    oracle.fill_vel_ghosts(A.g[L], ug)
    du = oracle.nodal_divergence(A.g[L], ug).ravel()
    P, offs = _interp(A)
    full = _split(A, P.T @ du, offs)
    ref = [np.where(A.dof[l], full[l] / 8.0 ** (L - l), 0.0) for l in range(L + 1)]
    bs = A.divergence(us)
    scale = max(np.abs(r).max() for r in ref)
    for l in range(L + 1):
        assert np.abs(bs[l] - ref[l]).max() < 1e-12 * scale, l


def test_two_levels_match_the_two_level_oracle(orc):
    gc = oracle.Geom.make((8, 8, 6), prob_hi=(1.0, 1.0, 0.75))
    boxes = [(4, 2, 0, 9, 7, 5), (10, 2, 0, 11, 5, 3)]
    A2 = AmrGeom(gc, boxes)
    A = AmrLevels(gc, [boxes])
    assert np.array_equal(A.dof[0], A2.cdof) and np.array_equal(A.ftype[1], A2.ftype)
    rng = np.random.default_rng(4)
    sig_c = 0.5 + rng.random(gc.cell_shape())
    sig_f = 0.5 + rng.random(A2.gf.cell_shape())
    xc = np.where(A2.cdof, rng.standard_normal(gc.node_shape()), 0.0)
    xf = np.where(A2.ftype == A2.DOF, rng.standard_normal(A2.gf.node_shape()), 0.0)
    yc, yf = A2.apply(sig_c, sig_f, xc, xf)
    ys = A.apply([sig_c, sig_f], [xc, xf])
    assert np.abs(ys[0] - yc).max() < 1e-13 * np.abs(yc).max()
    assert np.abs(ys[1] - yf).max() < 1e-13 * np.abs(yf).max()
    uc = rng.standard_normal((3,) + gc.cell_shape())
    uf = rng.standard_normal((3,) + A2.gf.cell_shape())
    rho = [1.0 + rng.random(gc.cell_shape()), 1.0 + rng.random(A2.gf.cell_shape())]
    sigs = [0.01 / r for r in rho]
    ucn, ufn, phc, phf, gpc, gpf, _ = A2.project(sigs[0], sigs[1], uc, uf, rtol=1e-13)
    un, ph, gp, _ = A.project(sigs, [uc, uf], rtol=1e-13)
    for a, b in ((un[0], ucn), (un[1], ufn), (ph[0], phc), (ph[1], phf), (gp[0], gpc), (gp[1], gpf)):
        assert np.abs(a - b).max() < 1e-10 * np.abs(b).max()


def test_symmetric_form_and_nullspace(orc):
    A = _levels(PER)
    sigs = _rand_sigs(A, 5)
    rng = np.random.default_rng(6)
    sc = [64.0, 8.0, 1.0]

    def rnd():
        return [np.where(A.dof[l], rng.standard_normal(A.g[l].node_shape()), 0.0) for l in range(3)]

    xs, zs = rnd(), rnd()
    ax, az = A.apply(sigs, xs), A.apply(sigs, zs)
    lhs = sum(sc[l] * np.sum(ax[l] * zs[l]) for l in range(3))
    rhs = sum(sc[l] * np.sum(az[l] * xs[l]) for l in range(3))
    mag = sum(sc[l] * np.abs(ax[l] * zs[l]).sum() for l in range(3))
    assert abs(lhs - rhs) < 1e-12 * mag
    ones = A.apply(sigs, [m.astype(float) for m in A.dof])
    assert max(np.abs(v).max() for v in ones) < 1e-10 * max(np.abs(v).max() for v in ax)


@pytest.mark.parametrize("case", [PER, BND], ids=["periodic", "bounded"])
def test_three_level_projection(orc, case):
    A = _levels(case)
    rng = np.random.default_rng(7)
    sigs = [0.01 / (1.0 + 0.5 * rng.random(g.cell_shape())) for g in A.g]
    us = [rng.standard_normal((3,) + g.cell_shape()) for g in A.g]
    b0 = A.divergence(us)
    un, ph, gp, its = A.project(sigs, us, rtol=1e-13)
    # phi solves the composite system (periodic: the weighted mean of b removed)
    ys = A.apply(sigs, [np.where(A.dof[l], ph[l], 0.0) for l in range(3)])
    sc = [64.0, 8.0, 1.0]
    mu = 0.0
    if A.singular():
        mu = sum(sc[l] * b0[l][A.dof[l]].sum() for l in range(3)) / sum(m.sum() for m in A.dof)
    bmax = max(np.abs(b).max() for b in b0)
    for l in range(3):
        r = np.where(A.dof[l], b0[l] - mu / sc[l] - ys[l], 0.0)
        assert np.abs(r).max() < 1e-11 * bmax, l
    # each level's cells outside the next level: its own gradient of its phi
    for l in range(3):
        g = oracle.nodal_gradient(A.g[l], ph[l])
        k = A.keep[l]
        assert np.allclose(un[l][:, k], (us[l] - sigs[l][None] * g)[:, k], rtol=0, atol=1e-14)
        assert np.allclose(gp[l][:, k], g[:, k], rtol=0, atol=1e-14)
    # covered cells hold the mean of their children, from the finest level down
    for l in (1, 0):
        c = A.cov[l]
        assert np.allclose(un[l][:, c], A.average_down(un[l + 1])[:, c], rtol=0, atol=1e-15)
        assert np.allclose(gp[l][:, c], A.average_down(gp[l + 1])[:, c], rtol=0, atol=1e-15)
    # phi under a finer level is the coincident finer value; dF from the coarser level
    for l in (1, 0):
        f = ph[l + 1][::2, ::2, ::2][:ph[l].shape[0], :ph[l].shape[1], :ph[l].shape[2]]
        assert np.array_equal(ph[l][A.allcov[l]], f[A.allcov[l]])
    for l in (1, 2):
        p = oracle.nodal_prolong_add(A.g[l], ph[l - 1], np.zeros(A.g[l].node_shape()))
        bd = (A.ftype[l] == A.BDRY) & ~A.dmask[l]
        assert np.abs(ph[l][bd] - p[bd]).max() < 1e-13 * np.abs(ph[l]).max()
    if A.singular():
        assert abs(ph[0].mean()) < 1e-14 * np.abs(ph[0]).max()
    else:
        assert not ph[0][A.dmask[0]].any()


def test_nesting_and_inflow_rules(orc):
    gc = oracle.Geom.make((8, 8, 8))
    with pytest.raises(AssertionError, match="properly nested"):
        AmrLevels(gc, [[(2, 2, 2, 11, 11, 11)], [(4, 4, 4, 11, 11, 11)]])
    gi = oracle.Geom.make((8, 8, 8), periodic=(0, 1, 1), bclo=(3, 0, 0), bchi=(2, 0, 0))
    with pytest.raises(AssertionError, match="inflow"):
        AmrLevels(gi, [[(0, 2, 2, 5, 7, 7)]])
