"""The FP32 Jacobi sweep variants selected per level through the
launch-shape table -- LDS-direct plane loads (nodal_jacobi_dma_kernel) and
wave strips of 2 or 4 rows with cross-lane shuffles
(nodal_jacobi_strip_kernel) -- compute every node with the same formula as
the register-staged sweep (the compiler may contract the
FP32 arithmetic differently, so the two agree to FP32 rounding, not bit for
bit).  Checked on a periodic cube and on a bounded, anisotropic box whose
tiles are partial in x and y: after one FP64 refinement step of three FP32
V-cycles (rtol 1e-2) the iterates agree to 1e-5 relative -- a wrong value on
any tile edge would show at the 1e-3 level -- and a full solve to 1e-10 takes
the same V-cycles either way."""
import re
import sys

import numpy as np
import pytest

from incflo_amd import boxes as B

pytestmark = pytest.mark.gpu

D, N = 1, 2


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from incflo_amd import mgx
    c = mgx.Context(0)
    yield c
    c.close()


def _force(table, dma):
    """every 4-byte Jacobi line of the table -> npt 2, 1024 workgroups, variant dma"""
    out = []
    for l in table.splitlines():
        f = l.split()
        if f and f[0] == "stencil" and f[4] == "4" and f[5] == "4" and f[6] == "0":
            f[8:12] = ["2", "1024", "2", str(dma)]
        out.append(" ".join(f))
    return "\n".join(out) + "\n"


@pytest.mark.parametrize("variant", [1, 2, 3])
@pytest.mark.parametrize("case", ["periodic", "bounded"])
def test_sweep_variants_match_the_staged_sweep(ctx, orc, case, variant):
    from incflo_amd import mgx
    if case == "periodic":
        n, per, bclo, bchi, ph = (64, 64, 64), (1, 1, 1), None, None, (1.0, 1.0, 1.0)
    else:
        n, per, bclo, bchi, ph = (40, 24, 20), (0, 0, 0), (D, N, N), (N, D, N), (2.0, 1.0, 1.5)
    _say(f"case sweep variant {variant} {case} n={n}")
    g = orc.Geom.make(n, periodic=per, bclo=bclo or (0, 0, 0), bchi=bchi or (0, 0, 0),
                      prob_hi=ph)
    rng = np.random.default_rng(3)
    sigma = 1.0 + 0.5 * np.tanh(rng.standard_normal(g.cell_shape()))
    b = rng.standard_normal(g.node_shape())
    bxs = B.chop_domain(n, 16)
    saved = mgx.tuning_get()
    P = mgx.NodalProjector(ctx, mgx.make_level(n, bxs, periodic=per, prob_hi=ph))
    if bclo is not None:
        P.setDomainBC(bclo, bchi)
    P.set_sigma(ctx.array(B.to_boxed(sigma, bxs, 0)), 0)
    rd = ctx.array(B.to_boxed(b, bxs, 0, nodal=True, periodic=per))

    def solve(rtol):
        xd = ctx.array(B.field_size(bxs, 0, 1, True))
        info = P.solve(xd, 0, rd, 0, rtol=rtol, mode=mgx.MODE_MIXED)
        return xd.download(), info

    try:
        solve(1e-10)
        rep = P.autotune(mgx.MODE_MIXED)  # fills the table for every level shape
        table = mgx.tuning_get()
        assert re.search(r"^stencil( \d+){3} 4 4 0 ", table, re.M), table
        out = {}
        for dma in (0, variant):
            mgx.tuning_set(_force(table, dma))
            out[dma] = (solve(1e-2), solve(1e-10))
        (a0, ia0), (x0, i0) = out[0]
        (a1, ia1), (x1, i1) = out[variant]
        assert ia0.vcycles == ia1.vcycles <= 3
        assert np.max(np.abs(a1 - a0)) <= 1e-5 * np.max(np.abs(a0))
        assert i0.converged and i1.converged and i1.vcycles == i0.vcycles
        assert np.max(np.abs(x1 - x0)) <= 1e-8 * np.max(np.abs(x0))
        _say("\n".join(l for l in rep.splitlines() if " 4 4 0 " in l))
    finally:
        mgx.tuning_set(saved)
        P.close()
