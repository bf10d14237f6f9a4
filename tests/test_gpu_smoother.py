"""The smoother choice of the nodal V-cycle (mgx_mlmg_opts.smoother): weighted
Jacobi (MLNodeLaplacian's GPU smoother, the FP64 default) and the Chebyshev
sweeps (the default of the FP32 V-cycles of the mixed and learned modes).
Both reach the oracle's FP64 solution (1e-10 relative in FP64, 1e-7 in the
mixed form, the bars of the nodal parity tests), and the Chebyshev factors
never need more V-cycles than Jacobi on the variable-density case.  The
factors themselves are pinned on the CPU side by the symbol of the operator
(tests/test_oracle_nodal.py covers the Jacobi scaling; the Chebyshev interval
is the same symbol's high-frequency range)."""
import sys

import numpy as np
import pytest

from incflo_amd import boxes as B

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from incflo_amd import mgx
    c = mgx.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_chebyshev_and_jacobi_reach_the_oracle(ctx, orc, mode):
    from incflo_amd import mgx, workloads
    n = 48
    _say(f"case smoother n={n} mode={mode}")
    rho, ustar, dt = workloads.variable_density_case(n, 3)
    g = orc.Geom.make((n, n, n))
    sigma = dt / rho
    rng = np.random.default_rng(5)
    b = rng.standard_normal(g.node_shape())
    b -= b.mean()
    xo, io, _ = orc.nodal_solve(g, sigma, b, rtol=1e-12)
    bxs = B.chop_domain((n, n, n), 16)
    P = mgx.NodalProjector(ctx, mgx.make_level((n, n, n), bxs))
    P.set_sigma(ctx.array(B.to_boxed(sigma, bxs, 0)), 0)
    rd = ctx.array(B.to_boxed(b, bxs, 0, nodal=True))
    cyc = {}
    for sm in (mgx.SMOOTHER_JACOBI, mgx.SMOOTHER_CHEBYSHEV):
        P.set_options(mgx.default_opts(smoother=sm))
        xd = ctx.array(B.field_size(bxs, 0, 1, True))
        info = P.solve(xd, 0, rd, 0, rtol=1e-12 if mode == 0 else 1e-10, mode=mode)
        assert info.converged
        x = B.from_boxed(xd.download(), bxs, 0, g.node_shape(), nodal=True)
        x, xr = x - x.mean(), xo - xo.mean()
        tol = 1e-10 if mode == 0 else 1e-7
        assert np.max(np.abs(x - xr)) <= tol * np.max(np.abs(xr)), sm
        cyc[sm] = info.vcycles
    _say(f"  V-cycles jacobi {cyc[mgx.SMOOTHER_JACOBI]} chebyshev {cyc[mgx.SMOOTHER_CHEBYSHEV]}")
    assert cyc[mgx.SMOOTHER_CHEBYSHEV] <= cyc[mgx.SMOOTHER_JACOBI]
    with pytest.raises(mgx.MgxError):
        P.set_options(mgx.default_opts(smoother=3))
    P.close()
