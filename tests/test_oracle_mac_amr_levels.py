"""Pins of the composite MAC oracle over several levels
(oracle.mac_amr.MacAmrLevels):
  * one fine level: the pinned two-level oracle (MacAmr), operator and
    projection, periodic and bounded;
  * a level 1 covering the whole domain: the two-level problem of levels 1
    and 2 alone (level 0 then holds no unknown);
  * three levels, fine levels on Dirichlet and Neumann faces: the operator
    is symmetric negative semidefinite with zero row sums (constant nullspace
    only when singular), phi = x has zero composite Laplacian across every
    coarse-fine x face of both interfaces (pins the 1.5 h distance and the
    reflux at each), and the projected velocity has zero composite divergence
    with the faces under each finer level its mean."""
import numpy as np
import pytest

from oracle import mac_amr as M

BCD, BCN = 1, 2


def _geom(orc, n, periodic=(1, 1, 1), bclo=(0, 0, 0), bchi=(0, 0, 0), ph=(1.0, 1.0, 1.0)):
    return orc.Geom.make(n, periodic=periodic, bclo=bclo, bchi=bchi, prob_hi=ph)


def _refine(orc, g, f=2):
    return orc.Geom.make([f * g.n[d] for d in range(3)], periodic=[g.periodic[d] for d in range(3)],
                         bclo=[g.bclo[d] for d in range(3)], bchi=[g.bchi[d] for d in range(3)],
                         prob_hi=[g.dx[d] * g.n[d] for d in range(3)])


def _mask(shape_xyz, boxes):
    F = np.zeros((shape_xyz[2], shape_xyz[1], shape_xyz[0]), bool)
    for lo, hi in boxes:
        F[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
    return F


def _beta(orc, g, rng=None):
    b = [np.ones(orc.face_shape(g, d)) if rng is None else 0.5 + rng.random(orc.face_shape(g, d))
         for d in range(3)]
    for d in range(3):
        if g.periodic[d]:
            ax = 2 - d
            sl0 = [slice(None)] * 3
            sln = [slice(None)] * 3
            sl0[ax] = 0
            sln[ax] = -1
            b[d][tuple(sln)] = b[d][tuple(sl0)]
    return b


def _faces(orc, g, rng, A, l):
    u = [rng.standard_normal(orc.face_shape(g, d)) for d in range(3)]
    A._sync_periodic(u, A.n[l])
    return u


@pytest.mark.parametrize("bounded", [False, True])
def test_one_fine_level_is_the_two_level_oracle(orc, bounded):
    rng = np.random.default_rng(1)
    per = (0, 1, 1) if bounded else (1, 1, 1)
    g = _geom(orc, (8, 6, 6), per, (BCD, 0, 0), (BCN, 0, 0), (1.0, 0.75, 0.75))
    gf = _refine(orc, g)
    F = _mask(gf.n, [((4, 2, 0), (9, 9, 5)), ((10, 0, 8), (13, 3, 11))])
    A2 = M.MacAmr(g, F)
    A = M.MacAmrLevels(g, [F])
    bc_, bf = _beta(orc, g, rng), _beta(orc, gf, rng)
    phic, phif = rng.standard_normal(g.cell_shape()), rng.standard_normal(gf.cell_shape())
    y2 = A2.apply(bc_, bf, phif, phic)
    y = A.apply([bc_, bf], [phic, phif])
    # MacAmr orders the fine unknowns first
    ref = np.concatenate([y2[A2.nfu:], y2[:A2.nfu]])
    assert np.max(np.abs(y - ref)) <= 1e-12 * np.max(np.abs(ref))
    uc, uf = _faces(orc, g, rng, A, 0), _faces(orc, gf, rng, A, 1)
    ucn, ufn, pf, pc, _, _ = A2.project(bc_, bf, uc, uf)
    un, ph, _, _ = A.project([bc_, bf], [uc, uf])
    for d in range(3):
        assert np.max(np.abs(un[0][d] - ucn[d])) <= 1e-10 * np.max(np.abs(ucn[d]))
        assert np.max(np.abs(un[1][d] - ufn[d])) <= 1e-10 * np.max(np.abs(ufn[d]))
    assert np.max(np.abs(ph[0] - pc)) <= 1e-10 * np.max(np.abs(pc))
    assert np.max(np.abs(ph[1] - pf)) <= 1e-10 * np.max(np.abs(pf))


def test_full_level_one_is_the_two_level_problem_above_it(orc):
    rng = np.random.default_rng(2)
    g = _geom(orc, (4, 4, 4))
    g1, g2 = _refine(orc, g), _refine(orc, g, 4)
    F1 = np.ones(g1.cell_shape(), bool)
    F2 = _mask(g2.n, [((4, 4, 2), (11, 9, 7))])
    A = M.MacAmrLevels(g, [F1, F2])
    A2 = M.MacAmr(g1, F2)
    assert A.unk[0].sum() == 0
    b0, b1, b2 = _beta(orc, g, rng), _beta(orc, g1, rng), _beta(orc, g2, rng)
    p1, p2 = rng.standard_normal(g1.cell_shape()), rng.standard_normal(g2.cell_shape())
    y = A.apply([b0, b1, b2], [np.zeros(g.cell_shape()), p1, p2])
    y2 = A2.apply(b1, b2, p2, p1)
    ref = np.concatenate([y2[A2.nfu:], y2[:A2.nfu]])
    assert np.max(np.abs(y - ref)) <= 1e-12 * np.max(np.abs(ref))


THREE = dict(n=(8, 8, 8), per=(0, 1, 1), bclo=(BCD, 0, 0), bchi=(BCN, 0, 0),
             F1=[((0, 2, 2), (9, 11, 11)), ((12, 4, 4), (15, 11, 9))],
             F2=[((0, 6, 6), (13, 17, 17)), ((26, 10, 10), (31, 17, 15))])


# fully periodic: both fine levels inside the domain
THREE_PER = dict(n=(8, 8, 8), per=(1, 1, 1), bclo=(0, 0, 0), bchi=(0, 0, 0),
                 F1=[((2, 2, 2), (11, 11, 11)), ((12, 4, 4), (15, 11, 9))],
                 F2=[((6, 6, 6), (19, 17, 17)), ((24, 10, 10), (29, 17, 15))])


def _three(orc, per=None):
    c = THREE if per is None or not all(per) else THREE_PER
    p = c["per"]
    g = _geom(orc, c["n"], p, c["bclo"], c["bchi"])
    g1, g2 = _refine(orc, g), _refine(orc, g, 4)
    A = M.MacAmrLevels(g, [_mask(g1.n, c["F1"]), _mask(g2.n, c["F2"])])
    return A, (g, g1, g2)


@pytest.mark.parametrize("per", [(0, 1, 1), (1, 1, 1)])
def test_three_levels_symmetric_semidefinite(orc, per):
    A, gs = _three(orc, per)
    rng = np.random.default_rng(3)
    K = A.matrix([_beta(orc, g, rng) for g in gs]).toarray()
    assert np.max(np.abs(K - K.T)) <= 1e-13 * np.max(np.abs(K))
    ev = np.linalg.eigvalsh(K)
    assert ev.max() <= 1e-10 * np.abs(ev).max()
    if A.singular:
        assert np.max(np.abs(K.sum(axis=1))) <= 1e-12 * np.max(np.abs(K))
        assert np.sum(np.abs(ev) <= 1e-10 * np.abs(ev).max()) == 1
    else:
        assert ev.max() < 0


def test_three_levels_linear_flux_exact(orc):
    # slabs spanning y and z: every coarse-fine face is an x face, where phi = x
    # has the exact two-point flux at both interfaces
    g = _geom(orc, (8, 4, 4), (0, 1, 1), (BCN, 0, 0), (BCN, 0, 0), (2.0, 1.0, 1.0))
    g1, g2 = _refine(orc, g), _refine(orc, g, 4)
    F1 = _mask(g1.n, [((4, 0, 0), (11, 7, 7))])
    F2 = _mask(g2.n, [((12, 0, 0), (19, 15, 15))])
    A = M.MacAmrLevels(g, [F1, F2])
    gs = (g, g1, g2)
    phis = [(np.arange(gg.n[0]) + 0.5) * gg.dx[0] * np.ones(gg.cell_shape()) for gg in gs]
    L = A.levels(A.apply([_beta(orc, gg) for gg in gs], phis))
    for l in (1, 2):
        assert np.max(np.abs(L[l][A.unk[l]])) <= 1e-11, l
    inner = np.zeros(g.cell_shape(), bool)
    inner[:, :, 1:-1] = True
    assert np.max(np.abs(L[0][inner & A.unk[0]])) <= 1e-11
    assert np.max(np.abs(L[0][:, :, 0])) > 0.1  # the Neumann faces do not vanish


@pytest.mark.parametrize("per", [(0, 1, 1), (1, 1, 1)])
def test_three_level_projection_divergence_free(orc, per):
    A, gs = _three(orc, per)
    rng = np.random.default_rng(5)
    betas = [_beta(orc, g, rng) for g in gs]
    us = [_faces(orc, g, rng, A, l) for l, g in enumerate(gs)]
    un, ph, x, div0 = A.project(betas, us)
    div = A.divergence(un)
    assert np.max(np.abs(div)) <= 1e-10 * np.max(np.abs(div0))
    avg = A.average_down(un)
    for l in range(3):
        for d in range(3):
            assert np.array_equal(avg[l][d], un[l][d])
    # phi under a finer level is the mean of its 8 cells
    for l in (1, 0):
        f = ph[l + 1]
        m = sum(f[a::2, b::2, c::2] for a in (0, 1) for b in (0, 1) for c in (0, 1)) / 8.0
        assert np.max(np.abs(ph[l][A.cov[l]] - m[A.cov[l]])) <= 1e-14 * np.max(np.abs(f))


def test_nesting_is_required(orc):
    g = _geom(orc, (8, 8, 8))
    g1, g2 = _refine(orc, g), _refine(orc, g, 4)
    F1 = _mask(g1.n, [((4, 4, 4), (11, 11, 11))])
    F2 = _mask(g2.n, [((8, 8, 8), (15, 15, 15))])  # touches level 1's boundary
    with pytest.raises(ValueError, match="nested"):
        M.MacAmrLevels(g, [F1, F2])
