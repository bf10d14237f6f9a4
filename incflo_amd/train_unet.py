"""Trains the learned coarse-level predictor of the nodal projection on the GPU.

    python -m incflo_amd.train_unet --flows 256 --n 128 --level 2 --out incflo_amd/data/unet_b16_l4.npy

Pipeline (all GPU work goes through libmgx.so):
  1. seeded variable-density flows at n^3 (rho = 1 + 0.5 tanh(32 Fourier
     modes), k^-5/3 solenoidal velocity, u* = u - dt (u.grad)u), generated on
     host worker processes: generated at n^3 (the 256^3 single-level case's
     generator) or, for a fraction --c3-fraction of the flows, generated at
     (2n)^3 and averaged down to n^3 (the base level of the two-level case);
  2. each flow is projected by the plain FP64 MLMG to 1e-10 relative
     residual; the (rhs, phi) pair is read out at multigrid level `level` as
     the network's normalised input and target (mgx_nodal_training_pair);
  3. the pairs are augmented with the 48 symmetries of the cube (axis
     permutations and reflections map periodic nodal solutions to solutions;
     cell-centred log(sigma) moves with the cells, nodes with the nodes);
  4. Adam on the mean-squared error with the network's own BF16-MFMA forward
     and backward kernels (mgx_unet_train_step), cosine learning-rate decay;
  5. FP32 master parameters are saved as a .npy file.
A held-out set of flows reports the validation loss every epoch.
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from incflo_amd import boxes as B  # noqa: E402
from incflo_amd import mgx, workloads  # noqa: E402

DEFAULT_WEIGHTS = os.path.join(HERE, "data", "unet_b16_l4.npy")


def _flow(args):
    """Worker: one seeded flow, already in the one-box device layout; style
    'c3': the (2n)^3 fields averaged down to n^3 before the advection step."""
    n, seed, style = args
    if style == "c3":
        h = 1.0 / n
        rho = workloads.average_down(workloads.density_field(2 * n, seed))
        u = workloads.average_down(workloads.spectral_velocity(2 * n, seed))
        umax = np.max(np.sqrt(np.sum(u * u, axis=0)))
        dt = 0.5 * h / max(umax, 1e-30)
        ustar = workloads.advected_velocity(u, dt, h)
    else:
        rho, ustar, dt = workloads.variable_density_case(n, seed)
    bxs = [((0, 0, 0), (n - 1, n - 1, n - 1))]
    return seed, B.to_boxed(rho, bxs, 0), B.to_boxed(ustar, bxs, 1), dt


def _log(*a):
    print(*a, flush=True)


class PairMaker:
    """Runs FP64 projections of n^3 flows and reads out training pairs."""

    def __init__(self, ctx, n, level):
        self.ctx, self.n, self.level = ctx, n, level
        self.bxs = [((0, 0, 0), (n - 1, n - 1, n - 1))]
        self.P = mgx.NodalProjector(ctx, mgx.make_level((n, n, n), self.bxs))
        m = n >> level
        self.m = m
        self.vel = ctx.array(B.field_size(self.bxs, 1, 3))
        self.rho = ctx.array(B.field_size(self.bxs, 0, 1))
        self.gp = ctx.array(B.field_size(self.bxs, 0, 3))
        self.p = ctx.array(B.field_size(self.bxs, 0, 1, True))
        self.fin = ctx.array(8 * m ** 3, np.float32)
        self.tgt = ctx.array(m ** 3, np.float32)

    def pair(self, rho_boxed, vel_boxed, dt):
        self.vel.upload(vel_boxed)
        self.rho.upload(rho_boxed)
        mgx.lib().mgx_memset_d(self.ctx.h, self.gp.ptr, 0, self.gp.nbytes)
        info = self.P.apply_nodal_projection(self.vel, 1, self.gp, 0, self.p, 0, dt,
                                             density=self.rho, rtol=1e-10, atol=0.0,
                                             mode=mgx.MODE_FP64)
        assert info.converged, info.as_dict()
        self.P.training_pair(self.level, self.fin, self.tgt)
        m = self.m
        return (self.fin.download().reshape(m, m, m, 8)[..., :2].copy(),
                self.tgt.download().reshape(m, m, m))

    def close(self):
        self.P.close()


def symmetry(x, t, perm, flips):
    """Apply a cube symmetry to features x (m,m,m,2: node rhs, cell log sigma)
    and target t (m,m,m).  Arrays are (z, y, x).  Reflecting a periodic nodal
    index i -> (-i) mod m is flip + roll by one; cell index i -> m-1-i is a
    plain flip."""
    b, c = x[..., 0], x[..., 1]
    b = np.transpose(b, perm)
    c = np.transpose(c, perm)
    t = np.transpose(t, perm)
    for ax in range(3):
        if flips[ax]:
            b = np.roll(np.flip(b, ax), 1, ax)
            t = np.roll(np.flip(t, ax), 1, ax)
            c = np.flip(c, ax)
    out = np.zeros(b.shape + (8,), np.float32)
    out[..., 0] = b
    out[..., 1] = c
    return out, np.ascontiguousarray(t, np.float32)


PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def make_pairs(pool, ctx, n, level, seeds, c3_fraction=0.0):
    maker = PairMaker(ctx, n, level)
    X, T = [], []
    t0 = time.time()
    # every flow whose running share of c3 flows falls behind the fraction is one
    jobs = [(n, s, "c3" if int((q + 1) * c3_fraction) > int(q * c3_fraction) else "c2")
            for q, s in enumerate(seeds)]
    for q, (seed, rho, ustar, dt) in enumerate(pool.imap(_flow, jobs)):
        x, t = maker.pair(rho, ustar, dt)
        X.append(x)
        T.append(t)
        if (q + 1) % 32 == 0:
            _log(f"  pairs {q + 1}/{len(seeds)}  {time.time() - t0:.1f}s")
    maker.close()
    return np.stack(X), np.stack(T)


def train(args):
    seeds = list(range(args.seed_base, args.seed_base + args.flows + args.val_flows))
    # host workers start before this process touches the GPU
    with mp.get_context("spawn").Pool(args.workers) as pool:
        ctx = mgx.Context(0)
        _log(f"generating {len(seeds)} pairs at {args.n}^3, level {args.level}")
        X, T = make_pairs(pool, ctx, args.n, args.level, seeds, args.c3_fraction)
    rng = np.random.default_rng(args.seed)
    Xv, Tv = X[args.flows:], T[args.flows:]
    X, T = X[:args.flows], T[:args.flows]
    m = X.shape[1]
    net = mgx.UNet(ctx, base=args.base, levels=4, seed=args.seed)
    if args.init and os.path.exists(args.init):
        net.set_params(np.load(args.init))
    bs = args.batch
    shape = (bs, m, m, m)
    xd = ctx.array(bs * m ** 3 * 8, np.float32)
    td = ctx.array(bs * m ** 3, np.float32)
    outd = ctx.array(bs * m ** 3, np.float32)
    steps_per_epoch = max(1, (len(X) * args.aug) // bs)
    total = steps_per_epoch * args.epochs
    history = []

    def val_loss():
        if len(Xv) == 0:
            return float("nan")
        ls = []
        for i in range(0, len(Xv) - bs + 1, bs):
            xb = np.zeros((bs, m, m, m, 8), np.float32)
            xb[..., :2] = Xv[i:i + bs]
            xd.upload(xb.ravel())
            net.forward(xd, outd, shape)
            y = outd.download().reshape(bs, m, m, m)
            ls.append(np.mean((y - Tv[i:i + bs]) ** 2) / np.mean(Tv[i:i + bs] ** 2))
        return float(np.mean(ls))

    _log(f"training: {len(X)} flows x {args.aug} symmetries, batch {bs}, {total} steps")
    step = 0
    t0 = time.time()
    for ep in range(args.epochs):
        losses = []
        for _ in range(steps_per_epoch):
            idx = rng.integers(0, len(X), size=bs)
            xb = np.zeros((bs, m, m, m, 8), np.float32)
            tb = np.zeros((bs, m, m, m), np.float32)
            for q, i in enumerate(idx):
                perm = PERMS[rng.integers(6)] if args.aug > 1 else (0, 1, 2)
                flips = rng.integers(0, 2, size=3) if args.aug > 1 else (0, 0, 0)
                xb[q], tb[q] = symmetry(X[i], T[i], perm, flips)
            xd.upload(xb.ravel())
            td.upload(tb.ravel())
            lr = args.lr * 0.5 * (1.0 + np.cos(np.pi * step / total))
            losses.append(net.train_step(xd, td, shape, float(max(lr, 1e-6))))
            step += 1
            if step % 200 == 0:
                _log(f"  step {step}/{total} mse {np.mean(losses[-200:]):.4e} "
                     f"{time.time() - t0:.1f}s")
        vl = val_loss()
        rec = {"epoch": ep, "train_mse": float(np.mean(losses)), "val_rel_mse": vl,
               "seconds": time.time() - t0}
        history.append(rec)
        _log(json.dumps(rec))
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, net.get_params())
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump({"n": args.n, "level": args.level, "flows": args.flows, "c3_fraction": args.c3_fraction,
                   "val_flows": args.val_flows, "base": args.base, "levels": 4,
                   "epochs": args.epochs, "batch": bs, "lr": args.lr, "seed_base": args.seed_base,
                   "history": history}, f, indent=1)
    _log(f"saved {args.out}")
    net.close()
    ctx.close()


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--level", type=int, default=2)
    ap.add_argument("--flows", type=int, default=256)
    ap.add_argument("--c3-fraction", type=float, default=0.0,
                    help="share of the flows generated at (2n)^3 and averaged down to n^3")
    ap.add_argument("--val-flows", type=int, default=16)
    ap.add_argument("--seed-base", type=int, default=1000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--base", type=int, default=16)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--aug", type=int, default=48)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--workers", type=int, default=12)
    ap.add_argument("--init", default="")
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    args = ap.parse_args()
    train(args)


if __name__ == "__main__":
    main()
