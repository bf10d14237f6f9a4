"""Error paths of the projection entry points on the GPU: a non-finite input
ends the solve unconverged instead of passing the tolerance, direction-
dependent (in/out) faces are refused, and the options are validated the same
way for every solver handle."""
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
def test_nan_velocity_is_not_converged(ctx, mode):
    from incflo_amd import mgx
    n = 32
    _say(f"case NaN velocity, nodal projection mode {mode}")
    bxs = B.chop_domain((n, n, n), 16)
    P = mgx.NodalProjector(ctx, mgx.make_level((n, n, n), bxs))
    P.set_const_sigma(1.0)
    rng = np.random.default_rng(5)
    u = rng.standard_normal((3, n, n, n))
    u[1, 3, 4, 5] = np.nan
    vel = ctx.array(B.to_boxed(u, bxs, 1))
    with pytest.raises(mgx.MgxError) as e:
        P.project(vel, 1, 1e-10, mode=mode)
    assert e.value.code == mgx.MGX_ERR_NOT_CONVERGED
    P.close()


def test_nan_mac_velocity_is_not_converged(ctx):
    from incflo_amd import mgx
    n = 32
    _say("case NaN face velocity, MAC projection")
    bxs = B.chop_domain((n, n, n), 16)
    M = mgx.MacProjector(ctx, mgx.make_level((n, n, n), bxs))
    M.updateBeta(1.0)
    rng = np.random.default_rng(6)
    um = []
    for d in range(3):
        a = rng.standard_normal(B.face_field_size(bxs, 0, d))
        if d == 2:
            a[77] = np.nan
        um.append(ctx.array(a))
    with pytest.raises(mgx.MgxError) as e:
        M.project(um[0], um[1], um[2], 0, 1e-10, 0.0)
    assert e.value.code == mgx.MGX_ERR_NOT_CONVERGED
    M.close()


def test_inout_faces_refused(ctx):
    from incflo_amd import mgx
    n = 16
    bxs = B.chop_domain((n, n, n), 16)
    lev = mgx.make_level((n, n, n), bxs, periodic=(0, 1, 1))
    P = mgx.NodalProjector(ctx, lev)
    with pytest.raises(mgx.MgxError) as e:
        P.setDomainBC((mgx.BC_INOUT, 0, 0), (mgx.BC_DIRICHLET, 0, 0))
    assert e.value.code == mgx.MGX_ERR_UNSUPPORTED
    P.close()
    M = mgx.MacProjector(ctx, lev)
    with pytest.raises(mgx.MgxError) as e:
        M.setDomainBC((mgx.BC_NEUMANN, 0, 0), (mgx.BC_INOUT, 0, 0))
    assert e.value.code == mgx.MGX_ERR_UNSUPPORTED
    M.close()


def test_options_validated_on_every_handle(ctx):
    from incflo_amd import mgx
    n = 16
    bxs = B.chop_domain((n, n, n), 8)
    lev = mgx.make_level((n, n, n), bxs)
    fine = mgx.make_level((2 * n, 2 * n, 2 * n), [((8, 8, 8), (15, 15, 15))])
    handles = [mgx.NodalProjector(ctx, lev), mgx.MacProjector(ctx, lev),
               mgx.AmrNodalProjector(ctx, [lev, fine]), mgx.AmrMacProjector(ctx, [lev, fine])]
    for h in handles:
        for bad in (dict(omega=2.5), dict(maxiter=0), dict(nu1=0, nu2=0),
                    dict(cycles_per_update=9)):
            with pytest.raises(mgx.MgxError) as e:
                h.set_options(mgx.default_opts(**bad))
            assert e.value.code == mgx.MGX_ERR_ARG, (type(h).__name__, bad)
        h.set_options(mgx.default_opts(cycles_per_update=0))
        h.close()
