"""Pins the two-level composite nodal oracle (oracle/amr.py).

AMReX's coarse-fine nodal synchronisation is not vendored, so the composite
problem is the conforming Q1 finite-element one and these checks are what
that definition fixes:
  * the composite operator and divergence equal the Galerkin forms
    P^T L_u P and P^T D_u u~ of the uniformly fine operator (coarse
    coefficient / velocity copied to the children outside F), assembled here
    independently with an explicit sparse interpolation matrix P;
  * no fine boxes -> the coarse operator; fine boxes everywhere -> the fine one;
  * the symmetric form is symmetric and annihilates constants;
  * the composite solve satisfies its equation, and the projection applies
    the pinned single-level gradient on each level with the average_down rule.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from oracle.amr import AmrGeom

BOXES = [(4, 2, 0, 9, 7, 5), (10, 2, 0, 11, 5, 3)]          # an L-shaped fine region
WRAP = [(0, 4, 4, 3, 9, 9), (12, 4, 4, 15, 9, 9)]            # crosses the periodic x boundary


def _geom(boxes, n=(8, 8, 6)):
    gc = oracle.Geom.make(n, prob_hi=(1.0, 1.0, 0.75))
    return AmrGeom(gc, boxes)


def _interp_matrix(A):
    """P: composite unknowns (coarse nodes, then fine nodes) -> every fine node"""
    gc, gf = A.gc, A.gf
    ncz, ncy, ncx = gc.node_shape()
    nfz, nfy, nfx = gf.node_shape()
    nc = ncz * ncy * ncx
    rows, cols, vals = [], [], []
    for k in range(nfz):
        for j in range(nfy):
            for i in range(nfx):
                r = i + nfx * (j + nfy * k)
                if A.ftype[k, j, i] == A.DOF:
                    rows.append(r); cols.append(nc + r); vals.append(1.0)
                    continue
                parts = []
                for f, nn in ((i, ncx), (j, ncy), (k, ncz)):
                    lo = f // 2
                    if f % 2 == 0:
                        parts.append([(lo % nn, 1.0)])
                    else:
                        parts.append([(lo % nn, 0.5), ((lo + 1) % nn, 0.5)])
                for (I, wx) in parts[0]:
                    for (J, wy) in parts[1]:
                        for (K, wz) in parts[2]:
                            assert A.cdof[K, J, I], "interpolation uses a covered coarse node"
                            rows.append(r); cols.append(I + ncx * (J + ncy * K))
                            vals.append(wx * wy * wz)
    return sp.csr_matrix((vals, (rows, cols)), shape=(nfz * nfy * nfx, nc + nfz * nfy * nfx))


def _uniform_sigma(A, sig_c, sig_f):
    rep = np.repeat(np.repeat(np.repeat(sig_c, 2, 0), 2, 1), 2, 2)
    return np.where(A.fmask, sig_f, rep)


def _fine_operator_matrix(A, sig_u):
    shape = A.gf.node_shape()
    n = int(np.prod(shape))
    cols = []
    for q in range(n):
        e = np.zeros(n)
        e[q] = 1.0
        cols.append(sp.csc_matrix(oracle.nodal_apply(A.gf, sig_u, e.reshape(shape)).reshape(n, 1)))
    return sp.hstack(cols).tocsr()


@pytest.mark.parametrize("boxes", [BOXES, WRAP])
def test_composite_operator_is_galerkin(orc, boxes):
    A = _geom(boxes)
    rng = np.random.default_rng(1)
    sig_c = 0.5 + rng.random(A.gc.cell_shape())
    sig_f = 0.5 + rng.random(A.gf.cell_shape())
    assert (A.ftype == A.DOF).sum() > 0 and (A.ftype == A.BDRY).sum() > 0 and (~A.cdof).sum() > 0
    P = _interp_matrix(A)
    Lu = _fine_operator_matrix(A, _uniform_sigma(A, sig_c, sig_f))
    xc = np.where(A.cdof, rng.standard_normal(A.gc.node_shape()), 0.0)
    xf = np.where(A.ftype == A.DOF, rng.standard_normal(A.gf.node_shape()), 0.0)
    x = np.concatenate([xc.ravel(), xf.ravel()])
    full = P.T @ (Lu @ (P @ x))
    nc = xc.size
    ref_c = np.where(A.cdof, full[:nc].reshape(xc.shape) / 8.0, 0.0)
    ref_f = np.where(A.ftype == A.DOF, full[nc:].reshape(xf.shape), 0.0)
    yc, yf = A.apply(sig_c, sig_f, xc, xf)
    scale = max(np.abs(ref_c).max(), np.abs(ref_f).max())
    assert np.abs(yc - ref_c).max() < 1e-12 * scale
    assert np.abs(yf - ref_f).max() < 1e-12 * scale


@pytest.mark.parametrize("boxes", [BOXES, WRAP])
def test_composite_divergence_is_galerkin(orc, boxes):
    A = _geom(boxes)
    rng = np.random.default_rng(2)
    uc = rng.standard_normal((3,) + A.gc.cell_shape())
    uf = rng.standard_normal((3,) + A.gf.cell_shape())
    rep = np.repeat(np.repeat(np.repeat(uc, 2, 1), 2, 2), 2, 3)
    uu = np.where(A.fmask[None], uf, rep)
    ug = np.zeros(A.gf.ghost_vec_shape())
    ug[:, 1:-1, 1:-1, 1:-1] = uu
    oracle.fill_vel_ghosts(A.gf, ug)
    du = oracle.nodal_divergence(A.gf, ug).ravel()
    P = _interp_matrix(A)
    full = P.T @ du
    nc = int(np.prod(A.gc.node_shape()))
    ref_c = np.where(A.cdof, full[:nc].reshape(A.gc.node_shape()) / 8.0, 0.0)
    ref_f = np.where(A.ftype == A.DOF, full[nc:].reshape(A.gf.node_shape()), 0.0)
    bc, bf = A.divergence(uc, uf)
    scale = max(np.abs(ref_c).max(), np.abs(ref_f).max())
    assert np.abs(bc - ref_c).max() < 1e-12 * scale
    assert np.abs(bf - ref_f).max() < 1e-12 * scale


def test_limits_reduce_to_single_level(orc):
    rng = np.random.default_rng(3)
    # no fine cells: the coarse operator
    A = _geom([])
    sig_c = 0.5 + rng.random(A.gc.cell_shape())
    sig_f = 0.5 + rng.random(A.gf.cell_shape())
    xc = rng.standard_normal(A.gc.node_shape())
    yc, yf = A.apply(sig_c, sig_f, xc, np.zeros(A.gf.node_shape()))
    assert np.abs(yc - oracle.nodal_apply(A.gc, sig_c, xc)).max() < 1e-13 * np.abs(yc).max()
    assert not yf.any()
    # fine cells everywhere: the fine operator, no coarse unknowns
    A = _geom([(0, 0, 0, 15, 15, 11)])
    xf = rng.standard_normal(A.gf.node_shape())
    yc, yf = A.apply(sig_c, sig_f, np.zeros(A.gc.node_shape()), xf)
    assert not A.cdof.any() and not yc.any()
    assert np.abs(yf - oracle.nodal_apply(A.gf, sig_f, xf)).max() < 1e-13 * np.abs(yf).max()


def test_symmetric_form_and_nullspace(orc):
    A = _geom(BOXES)
    rng = np.random.default_rng(4)
    sig_c = 0.5 + rng.random(A.gc.cell_shape())
    sig_f = 0.5 + rng.random(A.gf.cell_shape())
    dof_f = A.ftype == A.DOF

    def rnd():
        return (np.where(A.cdof, rng.standard_normal(A.gc.node_shape()), 0.0),
                np.where(dof_f, rng.standard_normal(A.gf.node_shape()), 0.0))

    (xc, xf), (zc, zf) = rnd(), rnd()
    ac, af = A.apply(sig_c, sig_f, xc, xf)
    bc, bf = A.apply(sig_c, sig_f, zc, zf)
    lhs = 8 * np.sum(ac * zc) + np.sum(af * zf)
    rhs = 8 * np.sum(bc * xc) + np.sum(bf * xf)
    assert abs(lhs - rhs) < 1e-12 * (8 * np.abs(ac * zc).sum() + np.abs(af * zf).sum())
    oc, of = A.apply(sig_c, sig_f, A.cdof.astype(float), dof_f.astype(float))
    assert max(np.abs(oc).max(), np.abs(of).max()) < 1e-10 * np.abs(ac).max()


def test_projection_solves_composite_system(orc):
    A = _geom(BOXES)
    rng = np.random.default_rng(5)
    rho_c = 1.0 + 0.5 * rng.random(A.gc.cell_shape())
    rho_f = 1.0 + 0.5 * rng.random(A.gf.cell_shape())
    dt = 0.01
    sig_c, sig_f = dt / rho_c, dt / rho_f
    uc = rng.standard_normal((3,) + A.gc.cell_shape())
    uf = rng.standard_normal((3,) + A.gf.cell_shape())
    b0c, b0f = A.divergence(uc, uf)
    ucn, ufn, phc, phf, gc, gf, its = A.project(sig_c, sig_f, uc, uf, rtol=1e-13)
    # phi solves the composite system (the mean of b is removed: periodic)
    yc, yf = A.apply(sig_c, sig_f, np.where(A.cdof, phc, 0.0), np.where(A.ftype == A.DOF, phf, 0.0))
    dof_f = A.ftype == A.DOF
    mu = (8 * b0c.sum() + b0f.sum()) / (A.cdof.sum() + dof_f.sum())
    rc = np.where(A.cdof, b0c - mu / 8 - yc, 0.0)
    rf = np.where(dof_f, b0f - mu - yf, 0.0)
    b0 = max(np.abs(b0c).max(), np.abs(b0f).max())
    assert max(np.abs(rc).max(), np.abs(rf).max()) < 1e-11 * b0
    # the update: fine cells by the fine gradient of the filled fine phi,
    # coarse cells outside F by the coarse gradient
    gref = oracle.nodal_gradient(A.gf, phf)
    assert np.allclose(ufn[:, A.fmask], (uf - sig_f[None] * gref)[:, A.fmask], rtol=0, atol=1e-14)
    gcr = oracle.nodal_gradient(A.gc, phc)
    out = ~A.cov
    assert np.allclose(ucn[:, out], (uc - sig_c[None] * gcr)[:, out], rtol=0, atol=1e-14)
    # covered coarse cells hold the average of their children
    assert np.allclose(ucn[:, A.cov], A.average_down(ufn)[:, A.cov], rtol=0, atol=1e-14)
    # coarse phi under F is the coincident fine value
    assert np.array_equal(phc[~A.cdof], phf[::2, ::2, ::2][~A.cdof])
    assert abs(phc.mean()) < 1e-14 * np.abs(phc).max()


def test_average_down_faces_injection_and_divergence():
    """average_down_faces returns an injected coarse field unchanged and makes
    the coarse divergence of every covered coarse cell the mean of its 8 fine
    cells' divergences (the fine faces inside a coarse cell cancel)."""
    from oracle import amr as A
    nc = (8, 8, 8)
    rng = np.random.default_rng(40)
    F = np.zeros((16, 16, 16), bool)
    F[4:12, 2:10, 6:14] = True  # fine cells (aligned with coarse cells)
    fshape = lambda n, d: tuple(n[2 - a] + (1 if a == 2 - d else 0) for a in range(3))
    nf = (16, 16, 16)
    fc = [rng.standard_normal(fshape(nc, d)) for d in range(3)]
    # injection: fine faces on coarse face planes take the coarse value
    ff = []
    for d in range(3):
        a = rng.standard_normal(fshape(nf, d))
        ax = 2 - d
        rep = np.repeat(np.repeat(fc[d], 2, axis=[x for x in range(3) if x != ax][0]),
                        2, axis=[x for x in range(3) if x != ax][1])
        sl = [slice(None)] * 3
        sl[ax] = slice(0, None, 2)
        a[tuple(sl)] = rep
        ff.append(a)
    out = A.average_down_faces(F, ff, fc)
    for d in range(3):
        assert np.array_equal(out[d], fc[d])
    # divergence: random fine faces
    ff = [rng.standard_normal(fshape(nf, d)) for d in range(3)]
    out = A.average_down_faces(F, ff, fc)
    divf = sum(np.diff(ff[d], axis=2 - d) * nf[d] for d in range(3))
    divc = sum(np.diff(out[d], axis=2 - d) * nc[d] for d in range(3))
    mean8 = sum(divf[a::2, b::2, c::2] for a in (0, 1) for b in (0, 1) for c in (0, 1)) / 8
    cov = F[0::2, 0::2, 0::2]
    assert cov.sum() == 4 * 4 * 4
    assert np.max(np.abs(divc - mean8)[cov]) <= 1e-12 * np.max(np.abs(mean8))
    # faces away from the fine region keep their values
    assert np.array_equal(out[0][0], fc[0][0])
