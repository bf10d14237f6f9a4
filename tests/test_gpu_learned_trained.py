"""The committed trained network (incflo_amd/data/unet_b16_l4.npy) in the
learned + mixed nodal projection, on seeded variable-density flows that are
not in its training set (training used seeds 1000-3031; these are 1-3).

For each case the same projector runs MODE_FP64, MODE_MIXED and MODE_LEARNED
on the same input.  The learned form must converge to the same relative
residual as the plain method (1e-10, measured again with the oracle's FP64
operator at 128^3), keep the network's step (fallback_used == 0) and give the
same phi as FP64 to 1e-7.  V-cycle counts and one-call times of all three
modes are printed, which is the ablation of the network (learned vs mixed).
"""
import json
import os
import sys
import time

import numpy as np
import pytest

from incflo_amd import boxes as B

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "incflo_amd", "data", "unet_b16_l4.npy")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from incflo_amd import mgx
    c = mgx.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def net(ctx):
    from incflo_amd import mgx
    meta = json.load(open(os.path.splitext(WEIGHTS)[0] + ".json"))
    u = mgx.UNet(ctx, base=meta["base"], levels=meta["levels"])
    u.set_params(np.load(WEIGHTS))
    yield u, meta
    u.close()


def _run(ctx, n, seed, net, meta):
    from incflo_amd import mgx, workloads
    rho, ustar, dt = workloads.variable_density_case(n, seed)
    bxs = B.chop_domain((n, n, n), 32)
    P = mgx.NodalProjector(ctx, mgx.make_level((n, n, n), bxs))
    # the level whose size matches the training level's (32^3 at 128^3)
    level = int(round(np.log2(n / (meta["n"] >> meta["level"]))))
    P.set_unet(net, level)
    vel0 = ctx.array(B.to_boxed(ustar, bxs, 1))
    rd = ctx.array(B.to_boxed(rho, bxs, 0))
    out = {}
    for name, mode in (("fp64", mgx.MODE_FP64), ("mixed", mgx.MODE_MIXED),
                       ("learned", mgx.MODE_LEARNED)):
        ts = []
        for rep in range(3):  # the first call captures the V-cycle graphs
            vel = ctx.array(vel0.n)
            vel.copy_from(vel0)
            gp = ctx.array(B.field_size(bxs, 0, 3))
            pn = ctx.array(B.field_size(bxs, 0, 1, True))
            ctx.sync()
            t0 = time.perf_counter()
            info = P.apply_nodal_projection(vel, 1, gp, 0, pn, 0, dt, density=rd, rtol=1e-10,
                                            atol=0.0, mode=mode)
            ctx.sync()
            ts.append(time.perf_counter() - t0)
        phi = B.from_boxed(pn.download(), bxs, 0, (n, n, n), nodal=True)
        out[name] = (info, phi, 1e3 * min(ts[1:]))
        _say(f"  {name:8s} n={n} seed={seed} vcycles={info.vcycles} fp64_updates={info.iters} "
             f"fallback={info.fallback_used} rel_resid={info.rel_resid:.2e} "
             f"ms={out[name][2]:.2f}")
    P.close()
    return rho, ustar, dt, out


@pytest.mark.parametrize("n,seed", [(128, 1), (128, 2), (256, 3)])
def test_trained_network_learned_mode(ctx, net, orc, n, seed):
    u, meta = net
    _say(f"case trained network n={n} seed={seed}")
    rho, ustar, dt, out = _run(ctx, n, seed, u, meta)
    i64, p64, _ = out["fp64"]
    il, pl, _ = out["learned"]
    assert i64.converged and il.converged
    assert il.rel_resid <= 1e-10
    # the trained network's step is accepted (no plain-MLMG fallback)
    assert il.fallback_used == 0
    assert np.max(np.abs(pl - p64)) <= 1e-7 * np.max(np.abs(p64))
    if n <= 128:
        # the plain FP64 operator of the oracle judges the learned answer
        g = orc.Geom.make((n, n, n))
        ug = np.zeros(g.ghost_vec_shape())
        ug[:, 1:-1, 1:-1, 1:-1] = ustar
        orc.fill_vel_ghosts(g, ug)
        b = orc.nodal_divergence(g, ug)
        b -= b.mean()
        r = orc.nodal_residual(g, dt / rho, b, pl)
        rel = np.max(np.abs(r)) / np.max(np.abs(b))
        _say(f"  oracle rel. residual of the learned phi {rel:.2e}")
        assert rel <= 1e-10


def test_learned_graph_replays_match_plain_launches(ctx, net):
    """The learned projection with its network forward and V-cycles replayed
    from HIP graphs (captured on the first call, replayed on the next two)
    gives bit for bit the phi and velocity of plain launches: the split-K
    convolutions sum their splits in a fixed order and reset their arrival
    counters themselves, so every replay starts from the same state."""
    from incflo_amd import mgx, workloads
    u, meta = net
    n, seed = 128, 2
    _say(f"case learned graphs vs plain launches n={n} seed={seed}")
    rho, ustar, dt = workloads.variable_density_case(n, seed)
    bxs = B.chop_domain((n, n, n), 32)
    P = mgx.NodalProjector(ctx, mgx.make_level((n, n, n), bxs))
    P.set_unet(u, int(round(np.log2(n / (meta["n"] >> meta["level"])))))
    vel0 = ctx.array(B.to_boxed(ustar, bxs, 1))
    rd = ctx.array(B.to_boxed(rho, bxs, 0))
    runs = []
    for graphs in (True, True, True, False):
        P.set_graphs(graphs)
        vel = ctx.array(vel0.n)
        vel.copy_from(vel0)
        gp = ctx.array(B.field_size(bxs, 0, 3))
        pn = ctx.array(B.field_size(bxs, 0, 1, True))
        info = P.apply_nodal_projection(vel, 1, gp, 0, pn, 0, dt, density=rd, rtol=1e-10, atol=0.0,
                                        mode=mgx.MODE_LEARNED)
        assert info.converged and info.fallback_used == 0
        runs.append((info.vcycles, pn.download(), vel.download()))
    P.close()
    for vc, ph, ve in runs[1:]:
        assert vc == runs[0][0]
        assert np.array_equal(ph, runs[0][1]) and np.array_equal(ve, runs[0][2])
