"""Pins the oracle's nodal solve with an overset mask (oracle.nodal_masked_solve),
the mechanism incflo uses for mixed boundaries in the nodal projection
(src/boundary_conditions/incflo_set_bcs.cpp:11-53: mask 0 = Dirichlet node):
  * a mask that is 0 exactly on a Dirichlet face reproduces the multigrid
    solve with that face Dirichlet (orc_nodal.c), independently computed;
  * the masked solve satisfies its equation on the free nodes and is 0 on the
    masked ones."""
import numpy as np

import oracle

D, N, INF = oracle.BC_DIRICHLET, oracle.BC_NEUMANN, oracle.BC_INFLOW


def test_mask_on_a_face_equals_dirichlet_face(orc):
    n = (16, 8, 8)
    rng = np.random.default_rng(1)
    gd = oracle.Geom.make(n, periodic=(0, 0, 1), bclo=(N, N, 0), bchi=(D, N, 0), prob_hi=(2, 1, 1))
    gn = oracle.Geom.make(n, periodic=(0, 0, 1), bclo=(N, N, 0), bchi=(N, N, 0), prob_hi=(2, 1, 1))
    sigma = 0.5 + rng.random(gd.cell_shape())
    b = rng.standard_normal(gd.node_shape())
    b[:, :, -1] = 0.0
    xd, info, ret = oracle.nodal_solve(gd, sigma, b, rtol=1e-13)
    assert ret == 0
    mask = np.ones(gn.node_shape(), bool)
    mask[:, :, -1] = False
    xm, its, ret2 = oracle.nodal_masked_solve(gn, sigma, b, mask, rtol=1e-13)
    assert ret2 == 0
    assert np.abs(xm - xd).max() < 1e-9 * np.abs(xd).max()


def test_partial_mask_solves_its_equation(orc):
    n = (16, 8, 8)
    rng = np.random.default_rng(2)
    g = oracle.Geom.make(n, periodic=(0, 0, 1), bclo=(INF, N, 0), bchi=(N, N, 0), prob_hi=(2, 1, 1))
    sigma = 0.5 + rng.random(g.cell_shape())
    b = rng.standard_normal(g.node_shape())
    mask = np.ones(g.node_shape(), bool)
    mask[:, :4, -1] = False          # lower half of the x-hi face fixed
    x, its, ret = oracle.nodal_masked_solve(g, sigma, b, mask, rtol=1e-13)
    assert ret == 0
    assert np.all(x[~mask] == 0.0)
    r = np.where(mask, b - oracle.nodal_apply(g, sigma, x), 0.0)
    assert np.abs(r).max() < 1e-12 * np.abs(b).max()
