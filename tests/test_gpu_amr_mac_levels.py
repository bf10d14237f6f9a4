"""GPU parity of the composite MAC projection over three AMR levels
(mgx_amr_mac_* with nlevels = 3) against the oracle
(oracle.mac_amr.MacAmrLevels, a sparse direct solve of the same composite
system), through the C ABI.

The GPU solves to 1e-12 relative residual (mixed FP32/FP64: to the same FP64
tolerance); the projected face velocities of every level that its boxes hold
(faces under a finer level averaged down) and phi on every level's cells are
compared at 1e-8 of their largest value.  Periodic, and x bounded with both
fine levels on the Dirichlet and the Neumann face."""
import sys

import numpy as np
import pytest

from incflo_amd import boxes as B
from oracle import mac_amr as M

pytestmark = pytest.mark.gpu

BCD, BCN = 1, 2


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from incflo_amd import mgx
    c = mgx.Context(0)
    yield c
    c.close()


PER = dict(per=(1, 1, 1), bclo=(0, 0, 0), bchi=(0, 0, 0),
           levels=[[((2, 2, 2), (11, 11, 11)), ((12, 4, 4), (15, 11, 9))],
                   [((6, 6, 6), (19, 17, 17)), ((24, 10, 10), (29, 17, 15))]])
BND = dict(per=(0, 1, 1), bclo=(BCD, 0, 0), bchi=(BCN, 0, 0),
           levels=[[((0, 2, 2), (9, 11, 11)), ((12, 4, 4), (15, 11, 9))],
                   [((0, 6, 6), (13, 17, 17)), ((26, 10, 10), (31, 17, 15))]])


def _mask(n, boxes):
    F = np.zeros((n[2], n[1], n[0]), bool)
    for lo, hi in boxes:
        F[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
    return F


@pytest.mark.parametrize("case", [PER, BND], ids=["periodic", "bounded"])
@pytest.mark.parametrize("mode", [0, 1], ids=["fp64", "mixed"])
def test_three_level_mac_matches_oracle(orc, ctx, case, mode):
    from incflo_amd import mgx
    nc = (8, 8, 8)
    per = case["per"]
    _say(f"case amr mac 3 levels {per} mode={mode}")
    gs = [orc.Geom.make([v << l for v in nc], periodic=per, bclo=case["bclo"], bchi=case["bchi"])
          for l in range(3)]
    Fs = [_mask(gs[l].n, case["levels"][l - 1]) for l in (1, 2)]
    A = M.MacAmrLevels(gs[0], Fs)
    rng = np.random.default_rng(11)
    betas = [A._sync_periodic([0.5 + rng.random(orc.face_shape(g, d)) for d in range(3)], A.n[l])
             for l, g in enumerate(gs)]
    us = [A._sync_periodic([rng.standard_normal(orc.face_shape(g, d)) for d in range(3)], A.n[l])
          for l, g in enumerate(gs)]
    un, ph, _, _ = A.project(betas, us)
    boxes = [B.chop_domain(nc, 4)] + case["levels"]
    levs = [mgx.make_level(tuple(gs[l].n), boxes[l], periodic=per) for l in range(3)]
    P = mgx.AmrMacProjector(ctx, levs)
    if not all(per):
        P.setDomainBC(case["bclo"], case["bchi"])
    P.initProjector([[ctx.array(B.faces_to_boxed(betas[l][d], boxes[l], 0, d)) for d in range(3)]
                     for l in range(3)], 0)
    ud = [[ctx.array(B.faces_to_boxed(us[l][d], boxes[l], 0, d)) for d in range(3)] for l in range(3)]
    phis = [ctx.array(B.field_size(boxes[l], 0, 1)) for l in range(3)]
    info = P.project(ud, 0, 1e-12, 0.0, phis=phis, mode=mode)
    _say(f"  iters {info.iters} rel {info.rel_resid:.3e}")
    assert info.converged and info.iters <= 40
    for l in range(3):
        g = gs[l]
        for d in range(3):
            got = B.faces_from_boxed(ud[l][d].download(), boxes[l], 0, d, orc.face_shape(g, d))
            held = B.faces_from_boxed(np.ones(B.face_field_size(boxes[l], 0, d)), boxes[l], 0, d,
                                      orc.face_shape(g, d)) > 0
            sc = np.max(np.abs(un[l][d][held]))
            err = np.max(np.abs(got[held] - un[l][d][held]))
            assert err <= 1e-8 * sc, (l, d, err, sc)
        gp = B.from_boxed(phis[l].download(), boxes[l], 0, g.cell_shape())
        m = A.F[l]
        sp = np.max(np.abs(ph[l][m]))
        assert np.max(np.abs(gp[m] - ph[l][m])) <= 1e-8 * sp, l
    P.close()


def test_mac_levels_must_nest(ctx):
    from incflo_amd import mgx
    nc = (8, 8, 8)
    cb = B.chop_domain(nc, 4)
    l1 = [((4, 4, 4), (11, 11, 11))]
    l2 = [((8, 8, 8), (15, 15, 15))]  # touches level 1's edge
    levs = [mgx.make_level(nc, cb), mgx.make_level((16, 16, 16), l1), mgx.make_level((32, 32, 32), l2)]
    P = mgx.AmrMacProjector(ctx, levs)
    beta = [[ctx.array(B.face_field_size(bx, 0, d)) for d in range(3)] for bx in (cb, l1, l2)]
    with pytest.raises(mgx.MgxError) as e:
        P.initProjector(beta, 0)
    assert e.value.code == mgx.MGX_ERR_ARG and "nested" in str(e.value)
    P.close()
