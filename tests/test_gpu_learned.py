"""GPU parity of the learned predictor's BF16-MFMA kernels against the oracle.

The oracle (oracle/orc_unet.c) runs in FP64 with the operands rounded to
bfloat16 where the matrix-core kernels round them, so the remaining
difference is FP32 accumulation.  Tolerances:
  * convolution forward / data gradient / weight gradient:
        |gpu - oracle| <= 2e-5 * M elementwise, M = the same sum over |terms|
        (an FP32 accumulation bound; an indexing or layout error is O(1) of M)
  * bias gradient (FP32 sums of FP32 values): 1e-5 relative
  * whole U-net: the GPU rounds FP32 activations to bfloat16 where the oracle
    rounds FP64 ones; the two differ in the last FP32 bits, so a few percent
    of activations round to the neighbouring bfloat16 value (2^-8 apart),
    about 6e-4 relative per layer, compounded over 14 layers (forward: 2e-2
    normwise).  In the backward pass the same noise flips ReLU masks of
    pre-activations near zero, each flip switching a whole voxel's
    contribution, so gradients are compared (a) in an affine regime
    (positive biases, small weights: no pre-activation near zero) at 1e-2
    normwise per parameter group, and (b) at the standard initialisation by
    cosine similarity >= 0.95 per group.  A wiring error (wrong skip channel,
    tap, pool index or routing) is O(1) in both.
Learned-mode projections must reach the same 1e-10 relative residual as the
other modes, whatever the network predicts (the plain MLMG judges it).
"""
import sys

import numpy as np
import pytest

from incflo_amd import boxes as B

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from incflo_amd import mgx
    c = mgx.Context(0)
    yield c
    c.close()


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CONV_CASES = [
    # (cin, cout, (batch, nz, ny, nx), periodic, relu)
    (8, 16, (1, 8, 8, 8), (1, 1, 1), True),
    (16, 32, (2, 4, 8, 16), (1, 0, 1), False),
    (48, 16, (1, 8, 8, 8), (0, 0, 0), True),
    (32, 64, (1, 4, 4, 8), (1, 1, 1), False),
    (64, 128, (1, 2, 4, 4), (1, 1, 1), True),
    (192, 64, (1, 4, 4, 4), (1, 1, 0), False),
    (16, 48, (1, 3, 5, 7), (1, 1, 1), False),  # ragged voxel count (105)
]


def _conv_data(cin, cout, shape, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(shape + (cin,))
    w = rng.standard_normal((cout, 27 * cin)) / np.sqrt(27 * cin)
    b = rng.standard_normal(cout)
    return x, w, b


@pytest.mark.parametrize("case", range(len(CONV_CASES)))
def test_conv_forward_matches_oracle(ctx, orc, case):
    from incflo_amd import mgx
    cin, cout, shape, per, relu = CONV_CASES[case]
    _say(f"case conv fwd cin={cin} cout={cout} shape={shape} per={per}")
    x, w, b = _conv_data(cin, cout, shape, case)
    xd = ctx.array(x.astype(np.float32).ravel(), np.float32)
    wd = ctx.array(w.astype(np.float32).ravel(), np.float32)
    bd = ctx.array(b.astype(np.float32), np.float32)
    od = ctx.array(int(np.prod(shape)) * cout, np.float32)
    mgx.conv3d_forward(ctx, xd, wd, bd, od, shape, cin, cout, periodic=per, relu=relu)
    got = od.download().reshape(shape + (cout,)).astype(np.float64)
    x32, w32, b32 = (a.astype(np.float32).astype(np.float64) for a in (x, w, b))
    ref = orc.conv3d_forward(x32, w32, b32, cout, periodic=per, relu=relu, bf16=True)
    mag = orc.conv3d_forward(np.abs(x32), np.abs(w32), np.abs(b32), cout, periodic=per, bf16=True)
    err = np.abs(got - ref)
    assert np.all(err <= 2e-5 * mag + 1e-30), float(np.max(err / (mag + 1e-30)))
    # the rounding points matter: without bf16 emulation the error is visible
    plain = orc.conv3d_forward(x32, w32, b32, cout, periodic=per, relu=relu, bf16=False)
    assert np.max(np.abs(got - plain)) > np.max(err)


@pytest.mark.parametrize("case", [1, 2, 3, 5, 6])
def test_conv_backward_matches_oracle(ctx, orc, case):
    from incflo_amd import mgx
    cin, cout, shape, per, _ = CONV_CASES[case]
    _say(f"case conv bwd cin={cin} cout={cout} shape={shape} per={per}")
    x, w, _b = _conv_data(cin, cout, shape, 100 + case)
    rng = np.random.default_rng(200 + case)
    dout = rng.standard_normal(shape + (cout,))
    nv = int(np.prod(shape))
    xd = ctx.array(x.astype(np.float32).ravel(), np.float32)
    wd = ctx.array(w.astype(np.float32).ravel(), np.float32)
    gd = ctx.array(dout.astype(np.float32).ravel(), np.float32)
    dind = ctx.array(nv * cin, np.float32)
    dwd = ctx.array(cout * 27 * cin, np.float32)
    dbd = ctx.array(cout, np.float32)
    mgx.conv3d_backward(ctx, xd, wd, gd, dind, dwd, dbd, shape, cin, cout, periodic=per)
    x32, w32, g32 = (a.astype(np.float32).astype(np.float64) for a in (x, w, dout))
    din, dw, db = orc.conv3d_backward(x32, w32, g32, periodic=per, bf16=True)
    mdin, mdw, _ = orc.conv3d_backward(np.abs(x32), np.abs(w32), np.abs(g32), periodic=per,
                                       bf16=True)
    got_din = dind.download().reshape(din.shape)
    got_dw = dwd.download().reshape(dw.shape)
    got_db = dbd.download()
    assert np.all(np.abs(got_din - din) <= 2e-5 * mdin + 1e-30)
    assert np.all(np.abs(got_dw - dw) <= 2e-5 * mdw + 1e-30)
    assert np.allclose(got_db, db, rtol=1e-5, atol=1e-5 * np.max(np.abs(db)))


def _unet_groups(base=16, cin0=8):
    C = [base << l for l in range(4)]
    convs = [(cin0, C[0]), (C[0], C[0]), (C[0], C[1]), (C[1], C[1]), (C[1], C[2]), (C[2], C[2]),
             (C[2], C[3]), (C[3], C[3]), (C[3] + C[2], C[2]), (C[2], C[2]),
             (C[2] + C[1], C[1]), (C[1], C[1]), (C[1] + C[0], C[0]), (C[0], C[0])]
    groups, off = [], 0
    for ci, co in convs:
        groups.append((off, off + co * 27 * ci))
        off += co * 27 * ci
        groups.append((off, off + co))
        off += co
    groups.append((off, off + base + 1))
    return groups


def _unet_input(shape, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros(shape + (8,))
    x[..., :2] = rng.standard_normal(shape + (2,))
    return x


@pytest.mark.parametrize("regime", ["affine", "relu"])
@pytest.mark.parametrize("per", [(1, 1, 1), (0, 1, 1)])
def test_unet_forward_and_gradients_match_oracle(ctx, orc, per, regime):
    from incflo_amd import mgx
    shape = (2, 16, 8, 16)
    _say(f"case unet gradients shape={shape} per={per} regime={regime}")
    net = mgx.UNet(ctx, base=16, levels=4, seed=7)
    p = net.get_params().astype(np.float64)
    assert p.size == orc.unet_nparam(16, 4, 8)
    rng = np.random.default_rng(3)
    groups = _unet_groups()
    for a, b in groups[1:-1:2]:  # conv biases
        p[a:b] = 1.0 if regime == "affine" else 0.05 * rng.standard_normal(b - a)
    if regime == "affine":
        for a, b in groups[0:-1:2]:  # conv weights
            p[a:b] *= 0.2
    net.set_params(p.astype(np.float32))
    p = p.astype(np.float32).astype(np.float64)
    x = _unet_input(shape, 1)
    xd = ctx.array(x.astype(np.float32).ravel(), np.float32)
    od = ctx.array(int(np.prod(shape)), np.float32)
    net.forward(xd, od, shape, periodic=per)
    x32 = x.astype(np.float32).astype(np.float64)
    ref = orc.unet_forward(p, x32, 16, 4, periodic=per, bf16=True)
    got = od.download().reshape(shape)
    ef = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    _say(f"  forward rel err {ef:.2e}")
    assert ef <= 2e-2
    dout = rng.standard_normal(shape)
    gd = ctx.array(dout.astype(np.float32).ravel(), np.float32)
    g = net.gradients(xd, gd, shape, periodic=per).astype(np.float64)
    gref = orc.unet_grad(p, x32, dout.astype(np.float32).astype(np.float64), 16, 4,
                         periodic=per, bf16=True)
    errs, coss = [], []
    for a, b in groups:
        ng = np.linalg.norm(gref[a:b])
        errs.append(np.linalg.norm(g[a:b] - gref[a:b]) / max(ng, 1e-30))
        coss.append(np.dot(g[a:b], gref[a:b]) / max(ng * np.linalg.norm(g[a:b]), 1e-30))
    _say("  gradient rel errs " + " ".join(f"{e:.1e}" for e in errs))
    _say("  gradient cosines " + " ".join(f"{c:.4f}" for c in coss))
    if regime == "affine":
        assert max(errs) <= 1e-2
    else:
        assert min(coss) >= 0.95
    net.close()


def test_unet_training_reduces_loss(ctx):
    from incflo_amd import mgx
    shape = (2, 16, 16, 16)
    _say("case unet training")
    net = mgx.UNet(ctx, base=16, levels=4, seed=1)
    x = _unet_input(shape, 5)
    t = 0.5 * x[..., 0] - 0.25 * x[..., 1]
    xd = ctx.array(x.astype(np.float32).ravel(), np.float32)
    td = ctx.array(t.astype(np.float32).ravel(), np.float32)
    losses = [net.train_step(xd, td, shape, 1e-3) for _ in range(120)]
    assert np.all(np.isfinite(losses))
    assert np.mean(losses[-10:]) < 0.5 * losses[0], (losses[0], losses[-10:])
    net.close()


def test_unet_training_is_reproducible(ctx):
    """Two networks from one seed trained on the same data follow the same
    trajectory bit for bit: every gradient and loss reduction sums its
    partials in a fixed order.  The second network first takes gradients of a
    smaller batch (parameters unchanged), so its workspaces are sized for
    that batch and then grown; the two gradients of one batch are equal too."""
    from incflo_amd import mgx
    shape, small = (2, 16, 16, 16), (1, 16, 16, 16)
    _say("case unet training reproducible")
    x = _unet_input(shape, 7)
    t = 0.5 * x[..., 0] - 0.25 * x[..., 1]
    xs = _unet_input(small, 8)
    ts = xs[..., 0].copy()
    xd = ctx.array(x.astype(np.float32).ravel(), np.float32)
    td = ctx.array(t.astype(np.float32).ravel(), np.float32)
    runs = []
    xsd = ctx.array(xs.astype(np.float32).ravel(), np.float32)
    tsd = ctx.array(ts.astype(np.float32).ravel(), np.float32)
    grads = []
    for warm in (False, True):
        net = mgx.UNet(ctx, base=16, levels=4, seed=3)
        if warm:
            grads.append(net.gradients(xsd, tsd, small))
        losses = [net.train_step(xd, td, shape, 1e-3) for _ in range(8)]
        runs.append((np.array(losses), net.get_params()))
        grads.append(net.gradients(xd, td, shape))
        net.close()
    assert np.array_equal(runs[0][0], runs[1][0]), (runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
    assert np.array_equal(grads[0], grads[2])
    assert np.all(np.isfinite(grads[1])) and runs[0][0][-1] < runs[0][0][0]


def _learned_case(ctx, n, seed, net, level, mode, opts=None):
    from incflo_amd import mgx, workloads
    rho, ustar, dt = workloads.variable_density_case(n, seed)
    bxs = B.chop_domain((n, n, n), 16)
    lev = mgx.make_level((n, n, n), bxs)
    P = mgx.NodalProjector(ctx, lev)
    if opts is not None:
        P.set_options(opts)
    if net is not None:
        P.set_unet(net, level)
    vel = ctx.array(B.to_boxed(ustar, bxs, 1))
    rd = ctx.array(B.to_boxed(rho, bxs, 0))
    gp = ctx.array(B.field_size(bxs, 0, 3))
    pn = ctx.array(B.field_size(bxs, 0, 1, True))
    info = P.apply_nodal_projection(vel, 1, gp, 0, pn, 0, dt, density=rd, rtol=1e-10,
                                    mode=mode)
    p = B.from_boxed(pn.download(), bxs, 0, (n, n, n), nodal=True)
    return P, info, p


def test_learned_mode_converges_and_falls_back(ctx, orc):
    from incflo_amd import mgx
    n = 32
    _say("case learned-mode projection n=32")
    net = mgx.UNet(ctx, base=16, levels=4, seed=3)
    # an untrained network with an amplified head: its prediction is garbage
    p = net.get_params()
    p[-17:-1] *= 1000.0
    net.set_params(p)
    P64, i64, p64 = _learned_case(ctx, n, 2, None, 0, mgx.MODE_FP64)
    PL, il, pl = _learned_case(ctx, n, 2, net, 2, mgx.MODE_LEARNED)
    assert il.converged and il.rel_resid <= 1e-10
    assert np.max(np.abs(pl - p64)) <= 1e-7 * np.max(np.abs(p64))
    # the garbage step raises the residual: it is rejected and plain MLMG runs
    assert il.fallback_used == 1
    P64.close()
    PL.close()
    net.close()


def test_learned_mode_errors(ctx):
    from incflo_amd import mgx
    n = 32
    bxs = B.chop_domain((n, n, n), 16)
    P = mgx.NodalProjector(ctx, mgx.make_level((n, n, n), bxs))
    P.set_const_sigma(1.0)
    rng = np.random.default_rng(0)
    vel = ctx.array(B.to_boxed(rng.standard_normal((3, n, n, n)), bxs, 1))
    with pytest.raises(mgx.MgxError) as e:  # no network attached
        P.project(vel, 1, 1e-10, mode=mgx.MODE_LEARNED)
    assert e.value.code == mgx.MGX_ERR_STATE
    net = mgx.UNet(ctx, base=16, levels=4, seed=3)
    P.set_unet(net, 3)  # 32^3 -> level 3 is 4^3: not a multiple of 8
    with pytest.raises(mgx.MgxError) as e:
        P.project(vel, 1, 1e-10, mode=mgx.MODE_LEARNED)
    assert e.value.code == mgx.MGX_ERR_UNSUPPORTED
    P.close()
    lev = mgx.make_level((n, n, n), bxs, periodic=(0, 1, 1))
    Q = mgx.NodalProjector(ctx, lev)
    Q.setDomainBC((2, 0, 0), (1, 0, 0))
    Q.set_const_sigma(1.0)
    Q.set_unet(net, 2)
    vel2 = ctx.array(B.to_boxed(rng.standard_normal((3, n, n, n)), bxs, 1))
    with pytest.raises(mgx.MgxError) as e:
        Q.project(vel2, 1, 1e-10, mode=mgx.MODE_LEARNED)
    assert e.value.code == mgx.MGX_ERR_UNSUPPORTED
    Q.close()
    net.close()


def test_training_pair_matches_oracle_features(ctx, orc):
    from incflo_amd import mgx
    n, level = 32, 2
    _say("case training pair n=32")
    P, info, p = _learned_case(ctx, n, 4, None, 0, mgx.MODE_FP64)
    m = n >> level
    xin = ctx.array(8 * m ** 3, np.float32)
    tgt = ctx.array(m ** 3, np.float32)
    P.training_pair(level, xin, tgt)
    f = xin.download().reshape(m, m, m, 8)
    t = tgt.download().reshape(m, m, m)
    assert np.all(f[..., 2:] == 0.0)
    assert abs(np.sqrt(np.mean(f[..., 0].astype(np.float64) ** 2)) - 1.0) < 1e-5
    assert abs(np.mean(f[..., 1].astype(np.float64))) < 1e-5
    # the target is the solution at every 4th node, scaled by one positive constant
    ps = p[::4, ::4, ::4]
    scale = np.sum(t * ps) / np.sum(t * t)
    assert scale > 0 and np.max(np.abs(t * scale - ps)) <= 1e-6 * np.max(np.abs(ps))
    P.close()
