"""AMReX-format output and restart from device fields (mgx_write_multifab,
mgx_read_multifab, mgx_write_plotfile, the checkpoint header), against
incflo_amd.amrex_io's reader and the downloaded fields: bit-exact round
trips, and a restarted run (incflo::ReadCheckpointFile,
src/utilities/io.cpp:113-239) that continues bit-identically to the
uninterrupted one."""
import os
import sys

import numpy as np
import pytest

from incflo_amd import amrex_io as AIO
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


def test_multifab_round_trip(ctx, tmp_path):
    from incflo_amd import mgx
    n = (24, 16, 20)
    _say("case multifab round trip")
    bxs = B.chop_domain(n, 12)
    lev = mgx.make_level(n, bxs)
    rng = np.random.default_rng(4)
    a = rng.standard_normal(B.field_size(bxs, 2, 3))
    d = ctx.array(a)
    pre = str(tmp_path / "Level_0" / "velocity")
    mgx.write_multifab(ctx, pre, lev, d, 2, 3)
    back = ctx.array(a.size)
    mgx.read_multifab(ctx, pre, lev, back, 2, 3)
    assert np.array_equal(back.download(), a)
    mf = AIO.read_multifab(pre)
    assert mf["ncomp"] == 3 and mf["ngrow"] == 2 and len(mf["boxes"]) == len(bxs)
    shape = (n[2], n[1], n[0])
    canon = B.from_boxed(a, bxs, 2, shape, ncomp=3)
    assert np.array_equal(AIO.to_canonical(mf, shape), canon)
    for b, (lo, hi) in enumerate(bxs):
        v = canon[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
        assert np.allclose(mf["min"][b], v.reshape(3, -1).min(axis=1), rtol=0, atol=0)
        assert np.allclose(mf["max"][b], v.reshape(3, -1).max(axis=1), rtol=0, atol=0)
    # nodal, no ghosts; a different box list is refused
    p = rng.standard_normal(B.field_size(bxs, 0, 1, True))
    pd = ctx.array(p)
    mgx.write_multifab(ctx, str(tmp_path / "p_nd"), lev, pd, 0, 1, nodal=True)
    pb = ctx.array(p.size)
    mgx.read_multifab(ctx, str(tmp_path / "p_nd"), lev, pb, 0, 1, nodal=True)
    assert np.array_equal(pb.download(), p)
    other = mgx.make_level(n, B.chop_domain(n, 8))
    with pytest.raises(mgx.MgxError):
        mgx.read_multifab(ctx, pre, other, ctx.array(B.field_size(B.chop_domain(n, 8), 2, 3)), 2, 3)


def test_plotfile_and_restart(ctx, tmp_path):
    from incflo_amd import mgx
    from incflo_amd.incflo import DeviceIncflo
    from oracle import incflo_step as S
    n, dt = 16, 0.016
    _say("case plotfile + checkpoint restart, Taylor vortex 16^3")
    dx = [2.0 / n] * 3
    v0 = S.taylor_vortex_velocity((n, n, n), dx)
    A = DeviceIncflo(ctx, (n, n, n), (2.0, 2.0, 2.0), 1e-3, v0, max_grid_size=8)
    A.initial_projection()
    A.initial_iterations(3, dt)
    for _ in range(4):
        A.advance(dt)
    chk = str(tmp_path / "chk00004")
    A.write_checkpoint(chk)
    plt = str(tmp_path / "plt00004")
    A.write_plotfile(plt)
    pf = AIO.read_plotfile(plt)
    assert pf["names"] == ["velx", "vely", "velz", "gpx", "gpy", "gpz", "density"]
    assert abs(pf["time"] - A.time) == 0.0 and pf["finest_level"] == 0
    assert np.allclose(pf["dx"][0], dx) and pf["prob_hi"] == [2.0, 2.0, 2.0]
    data = AIO.to_canonical(pf["levels"][0], (n, n, n))
    assert np.array_equal(data[0:3], A.velocity())
    for _ in range(4):
        A.advance(dt)
    # a fresh run restarted from the checkpoint continues identically
    Bd = DeviceIncflo(ctx, (n, n, n), (2.0, 2.0, 2.0), 1e-3, np.zeros_like(v0), max_grid_size=8)
    Bd.restart(chk)
    assert Bd.nstep == 4 and Bd.time == 4 * dt
    for _ in range(4):
        Bd.advance(dt)
    assert np.array_equal(Bd.velocity(), A.velocity())
    assert np.array_equal(Bd.pressure(), A.pressure())
    A.close()
    Bd.close()
