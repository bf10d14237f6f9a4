#!/usr/bin/env python3
"""Headline benchmark: nodal projections/s at 256^3 to 1e-10 relative residual.

One step = one incflo nodal projection (incflo::ApplyNodalProjection,
src/projection/incflo_apply_nodal_projection.cpp:29-267) through the C ABI
(mgx_apply_nodal_projection): gp add-back, sigma = dt/rho, rhs = D.u, MLMG to
||b - L phi||_inf / ||b||_inf <= 1e-10, u -= sigma G(phi), gp = G(phi),
p_nd = phi.  Before each step the projected velocity is restored from a
pristine device copy with the library's streaming copy kernel (about 0.15 ms);
the host-clock rate (`value`) includes that copy, the HIP-event times
(`ms_event_median`, taken on the solver stream around the projection call
alone) do not.

Modes: fp64 (plain FP64 MLMG, the denominator), mixed (FP32 V-cycles inside
FP64 iterative refinement) and learned (mixed plus the learned first step,
when trained weights are present).  The headline `value` is the learned +
mixed form (mixed without weights).

Seed 1 is the timed headline case (warm-up W, then K steps).  Seeds 2..5 run
every mode on their own flows (1 warm-up + 5 steps: V-cycles, residual and
the median event time).  Checks at seed 1: every mode's phi against the
oracle's FP64 operator (relative residual) and against an FP64 oracle solve
of the same problem (u and mean-free phi errors).  The CPU baseline is the
oracle's FP64 projection on the host, median of 5 after 1 warm-up, at a
bounded size (default 96^3), scaled to 256^3 by node count.

Prints ONE JSON line.  Only one GPU is used (--gpus > 1 exits).
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "nodal projections/s at 256^3, rel. resid 1e-10: learned+mixed vs plain FP64"
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Case:
    """Device buffers of one seeded 256^3 flow and its projector."""

    def __init__(self, ctx, n, seed, mgs, cache=False):
        from incflo_amd import boxes as B
        from incflo_amd import workloads as W
        path = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"mgx_case_{n}_{seed}.npz")
        if cache and os.path.exists(path):
            z = np.load(path)
            self.rho, self.ustar, self.dt = z["rho"], z["ustar"], float(z["dt"])
        else:
            log(f"bench: generating {n}^3 variable-density case seed={seed}")
            self.rho, self.ustar, self.dt = W.variable_density_case(n, seed)
            if cache:
                np.savez(path, rho=self.rho, ustar=self.ustar, dt=self.dt)
        self.n, self.seed = n, seed
        self.bxs = B.chop_domain((n, n, n), mgs)
        self.vel0 = ctx.array(B.to_boxed(self.ustar, self.bxs, 1))
        self.vel = ctx.array(self.vel0.n)
        self.rhod = ctx.array(B.to_boxed(self.rho, self.bxs, 0))
        self.gp = ctx.array(B.field_size(self.bxs, 0, 3))
        self.p = ctx.array(B.field_size(self.bxs, 0, 1, True))

    def phi(self):
        from incflo_amd import boxes as B
        n = self.n
        return B.from_boxed(self.p.download(), self.bxs, 0, (n, n, n), nodal=True)

    def velocity(self):
        from incflo_amd import boxes as B
        n = self.n
        return B.from_boxed(self.vel.download(), self.bxs, 1, (n, n, n), ncomp=3)

    def free(self):
        for a in (self.vel0, self.vel, self.rhod, self.gp, self.p):
            a.free()


def run_mode(ctx, P, case, mode, steps, warmup, rtol, timed_repeat=True, tune=False):
    """warmup + steps projections; per-step HIP-event times of the projection
    call, the host-clock total of the steps (restore copy included), and an
    instrumented repeat that times the finest-level sweeps.  With tune, the
    launch shapes are autotuned after the warm-up (then one more warm-up
    captures the V-cycle graphs with the tuned shapes)."""
    from incflo_amd import mgx
    c = case
    use_stream_copy = c.vel.nbytes % 16 == 0

    def restore():
        if use_stream_copy:  # the library's streaming copy kernel
            ctx.check(mgx.lib().mgx_stream_copy(ctx.h, c.vel.ptr, c.vel0.ptr, c.vel.nbytes))
        else:
            c.vel.copy_from(c.vel0)
        ctx.check(mgx.lib().mgx_memset_d(ctx.h, c.gp.ptr, 0, c.gp.nbytes))

    def project():
        return P.apply_nodal_projection(c.vel, 1, c.gp, 0, c.p, 0, c.dt, density=c.rhod,
                                        rho_ng=0, rtol=rtol, atol=0.0, mode=mode)

    for _ in range(warmup):
        restore()
        project()
    report = None
    if tune:
        report = P.autotune(mode)
        restore()
        project()
    ctx.sync()
    infos, ev = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        restore()
        ctx.record(0)
        infos.append(project())
        ctx.record(1)
        ev.append(ctx.elapsed_ms(0, 1))
    ctx.sync()
    t1 = time.perf_counter()
    ms, cnt = 0.0, 0
    if timed_repeat:
        # the finest-level smoother's launch durations come from HIP events on
        # the solver's stream in an instrumented repeat of the same steps
        # (events cannot sit inside a replayed graph, so these cycles run as
        # plain launches of the same kernels)
        P.set_timing(True)
        for _ in range(steps):
            restore()
            project()
        ctx.sync()
        ms, cnt = P.get_timing()
        P.set_timing(False)
    return (t1 - t0), infos, ev, ms, cnt, report


def measured_traffic(kernel_key, n):
    """HBM bytes per launch of the finest-level smoother from the committed
    counter-run summary (profiles/jacobi_256_traffic.txt, 256^3 only)."""
    path = os.path.join(ROOT, "profiles", "jacobi_256_traffic.txt")
    if n != 256 or not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes" and f[1] == kernel_key:
            return float(f[2])
    return None


def host_cpu():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(n_full, n, seed, threads):
    """Oracle FP64 nodal projection on the host at n^3 (same generator):
    median of 5 after 1 warm-up, rate scaled to n_full^3 by node count."""
    import oracle as orc
    orc.load()
    from incflo_amd import workloads as W
    rho, ustar, dt = W.variable_density_case(n, seed)
    g = orc.Geom.make((n, n, n))
    times, iters = [], 0
    for rep in range(6):
        ug = np.zeros(g.ghost_vec_shape())
        ug[:, 1:-1, 1:-1, 1:-1] = ustar
        orc.fill_vel_ghosts(g, ug)
        t0 = time.perf_counter()
        phi, gphi, info, ret = orc.nodal_project(g, dt / rho, ug, rtol=1e-10)
        if rep > 0:
            times.append(time.perf_counter() - t0)
        iters = info.iters
    t = statistics.median(times)
    scale = (n / n_full) ** 3
    model, ncpu = host_cpu()
    return {"value": scale / t, "unit": "projections/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle nodal projection at {n}^3 (same generator, seed {seed}, "
                      f"{iters} V-cycles), median of 5 after 1 warm-up = {t:.3f} s, "
                      f"rate scaled by ({n}/{n_full})^3",
            "host": {"cpu_model": model, "nproc": ncpu, "threads": threads},
            "median_s": t, "iters": iters}


def oracle_check(case, phis, vels):
    """FP64 residual of every mode's phi under the oracle operator, and the
    u / mean-free phi errors against an FP64 oracle solve."""
    import oracle as orc
    orc.load()
    n = case.n
    g = orc.Geom.make((n, n, n))
    ug = np.zeros(g.ghost_vec_shape())
    ug[:, 1:-1, 1:-1, 1:-1] = case.ustar
    orc.fill_vel_ghosts(g, ug)
    b = orc.nodal_divergence(g, ug)
    b -= b.mean()
    bn = np.max(np.abs(b))
    sigma = case.dt / case.rho
    out = {}
    for name, phi in phis.items():
        r = orc.nodal_residual(g, sigma, b, phi)
        out[name] = {"oracle_rel_resid": float(np.max(np.abs(r)) / bn)}
    log("bench: FP64 oracle solve at full size")
    t0 = time.perf_counter()
    phio, gphio, info, ret = orc.nodal_project(g, sigma, ug, rtol=1e-12)
    uo = ug[:, 1:-1, 1:-1, 1:-1]
    po = phio - phio.mean()
    for name, phi in phis.items():
        pm = phi - phi.mean()
        out[name]["phi_err_vs_oracle"] = float(np.max(np.abs(pm - po)) / np.max(np.abs(po)))
        out[name]["u_err_vs_oracle"] = float(np.max(np.abs(vels[name] - uo)) / np.max(np.abs(uo)))
    out["oracle_solve"] = {"rtol": 1e-12, "iters": info.iters, "seconds": time.perf_counter() - t0}
    return out


def tgv_workload(args):
    """The Taylor-Green vortex on 64^3 cells (unit cube, rho = 1, mu = 1e-3,
    max_grid_size 32): K incflo steps of the device predictor (Godunov
    extrapolation, MAC projection, convective term, implicit viscous solves,
    nodal projection; MLMG to 1e-10) after W untimed ones, each step timed
    with HIP events on the context's stream.  A secondary line, not the
    headline metric."""
    from incflo_amd import mgx
    from incflo_amd.incflo import DeviceIncflo
    n = args.n if args.n != 256 else 64
    h = 1.0 / n
    c = (np.arange(n) + 0.5) * h
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    tp = 2 * np.pi
    vel = np.stack([np.sin(tp * x) * np.cos(tp * y) * np.cos(tp * z),
                    -np.cos(tp * x) * np.sin(tp * y) * np.cos(tp * z), np.zeros_like(x)])
    ctx = mgx.Context(0)
    I = DeviceIncflo(ctx, (n, n, n), (1.0, 1.0, 1.0), 1e-3, vel, max_grid_size=32,
                     mac_rtol=1e-10, nodal_rtol=1e-10, diff_rtol=1e-10)
    dt = 0.5 * h  # CFL 0.5 at max |u| = 1
    I.initial_projection()
    for _ in range(args.warmup):
        I.advance(dt)
    ctx.sync()
    ev = []
    solves0 = dict(I.solves)
    for _ in range(args.steps):
        ctx.record(0)
        I.advance(dt)
        ctx.record(1)
        ev.append(ctx.elapsed_ms(0, 1))
    ctx.sync()
    med = statistics.median(ev)
    per = {k: (I.solves[k] - solves0[k]) / args.steps for k in I.solves}
    print(json.dumps({"metric": f"incflo steps/s at {n}^3 (Taylor-Green, device predictor)",
                      "value": 1e3 / med, "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
                      "warmup": args.warmup, "ms_per_step": med, "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                      "data": "synthetic", "config": {"workload": f"taylor-green {n}^3, rho 1, "
                                                      f"mu 1e-3, max_grid_size 32, dt {dt:g}",
                                                      "solves_per_step": per},
                      "ms_event_median": med, "ms_event_min": min(ev)}), flush=True)
    I.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="projection", choices=("projection", "tgv"),
                    help="projection: the 256^3 headline; tgv: the 64^3 Taylor-Green step")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--max-grid-size", type=int, default=32)
    ap.add_argument("--seeds", default="1,2,3,4,5",
                    help="the first is the timed headline case, the rest run 5 steps each")
    ap.add_argument("--modes", default="fp64,mixed,learned",
                    help="learned runs only when trained weights are present")
    ap.add_argument("--weights", default=os.path.join(ROOT, "incflo_amd", "data", "unet_b16_l4.npy"))
    ap.add_argument("--net-level", type=int, default=-1,
                    help="multigrid level the network acts on (default: the level whose size "
                         "matches the training level's)")
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--vcycle-bits", type=int, default=32, choices=(16, 32),
                    help="storage of the mixed forms' finest V-cycle level")
    ap.add_argument("--cycles-per-update", type=int, default=0,
                    help="FP32 V-cycles per FP64 refinement step in the mixed forms "
                         "(0: the library default)")
    ap.add_argument("--learned-steps", type=int, default=0,
                    help="learned form: refinement steps started from the network's "
                         "correction (0: the library default, the first; -1: every step)")
    ap.add_argument("--cpu-n", type=int, default=96, help="size of the CPU-baseline sample")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shapes (no mgx_nodal_autotune)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-oracle-solve", action="store_true")
    ap.add_argument("--cache", action="store_true",
                    help="keep the generated inputs in $TMPDIR between runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (got --gpus {args.gpus}); the multi-GPU "
              "sharding of this path is a design only", file=sys.stderr)
        sys.exit(2)

    if args.workload == "tgv":
        tgv_workload(args)
        return

    from incflo_amd import mgx

    n = args.n
    seeds = [int(s) for s in args.seeds.split(",") if s]
    ctx = mgx.Context(0)
    case = Case(ctx, n, seeds[0], args.max_grid_size, args.cache)
    P = mgx.NodalProjector(ctx, mgx.make_level((n, n, n), case.bxs))
    P.set_vcycle_bits(args.vcycle_bits)
    if args.cycles_per_update or args.learned_steps:
        P.set_options(mgx.default_opts(cycles_per_update=args.cycles_per_update,
                                       learned_steps=args.learned_steps))
    nnodes = n ** 3

    modes = {"fp64": mgx.MODE_FP64, "mixed": mgx.MODE_MIXED, "learned": mgx.MODE_LEARNED}
    results, phis, vels, tuned = {}, {}, {}, {}
    net = None
    mode_names = args.modes.split(",")
    if "learned" in mode_names:
        if os.path.exists(args.weights):
            meta_path = os.path.splitext(args.weights)[0] + ".json"
            meta = json.load(open(meta_path)) if os.path.exists(meta_path) else {}
            level = args.net_level
            if level < 0:
                m_train = meta.get("n", 128) >> meta.get("level", 2)
                level = int(round(np.log2(n / m_train)))
            net = mgx.UNet(ctx, base=meta.get("base", 16), levels=4)
            net.set_params(np.load(args.weights))
            P.set_unet(net, level)
            log(f"bench: learned predictor {args.weights} at level {level} ({n >> level}^3)")
        else:
            log(f"bench: no trained weights at {args.weights}; learned mode skipped")
            mode_names.remove("learned")
    for name in mode_names:
        mode = modes[name]
        log(f"bench: mode {name}: warmup {args.warmup}, steps {args.steps}")
        t, infos, ev, ms, cnt, rep = run_mode(ctx, P, case, mode, args.steps, args.warmup,
                                              args.rtol, tune=not args.no_tune)
        if rep:
            tuned[name] = rep.splitlines()
        # stored word of the finest-level V-cycle arrays (x, b, sigma)
        word = 8 if mode == mgx.MODE_FP64 else args.vcycle_bits // 8
        avg_ms = ms / max(cnt, 1)
        results[name] = {
            "projections_per_s": args.steps / t,
            "ms_per_projection": 1e3 * t / args.steps,
            "ms_event_median": statistics.median(ev),
            "ms_event_min": min(ev),
            "vcycles": infos[-1].vcycles,
            "fp64_updates": infos[-1].iters,
            "fallback_used": infos[-1].fallback_used,
            "rel_resid": infos[-1].rel_resid,
            "converged": all(i.converged for i in infos),
            "fine_smoother_avg_us": 1e3 * avg_ms,
            "fine_smoother_launches": cnt,
            "fine_smoother_GBps": 4 * word * nnodes / (avg_ms * 1e-3) / 1e9 if cnt else None,
        }
        log(f"bench: {name}: {results[name]}")
        if not args.no_check:
            phis[name] = case.phi()
            vels[name] = case.velocity()

    check = {}
    if phis and not args.no_check:
        if args.no_oracle_solve:
            import oracle as orc  # residual check only
            orc.load()
            g = orc.Geom.make((n, n, n))
            ug = np.zeros(g.ghost_vec_shape())
            ug[:, 1:-1, 1:-1, 1:-1] = case.ustar
            orc.fill_vel_ghosts(g, ug)
            b = orc.nodal_divergence(g, ug)
            b -= b.mean()
            for name, phi in phis.items():
                r = orc.nodal_residual(g, case.dt / case.rho, b, phi)
                check[name] = {"oracle_rel_resid": float(np.max(np.abs(r)) / np.max(np.abs(b)))}
        else:
            check = oracle_check(case, phis, vels)
        for other in ("mixed", "learned"):
            if "fp64" in phis and other in phis:
                check[f"{other}_vs_fp64_phi_rel"] = float(
                    np.max(np.abs(phis[other] - phis["fp64"])) / np.max(np.abs(phis["fp64"])))
        log(f"bench: oracle check {check}")
    phis.clear()
    vels.clear()

    # seeds 2..5: every mode on its own flow
    per_seed = {}
    for seed in seeds[1:]:
        case.free()
        case = Case(ctx, n, seed, args.max_grid_size, args.cache)
        per_seed[seed] = {}
        for name in mode_names:
            t, infos, ev, _, _, _ = run_mode(ctx, P, case, modes[name], 5, 1, args.rtol,
                                             timed_repeat=False)
            per_seed[seed][name] = {"ms_event_median": statistics.median(ev),
                                    "vcycles": infos[-1].vcycles,
                                    "fallback_used": infos[-1].fallback_used,
                                    "rel_resid": infos[-1].rel_resid,
                                    "converged": all(i.converged for i in infos)}
        log(f"bench: seed {seed}: {per_seed[seed]}")
    case.free()

    conv = {k: v for k, v in results.items() if v["converged"]}
    # the headline is the learned + mixed-precision form; without trained
    # weights the mixed form stands in (named in config.mode)
    head = next((k for k in ("learned", "mixed", "fp64") if k in conv), list(results)[0])
    hr = results[head]
    word = 8 if head == "fp64" else args.vcycle_bits // 8
    wname = {8: "FP64", 4: "FP32", 2: "FP16-stored, FP32 arithmetic"}[word]
    roofline = None
    if hr["fine_smoother_launches"]:
        achieved = hr["fine_smoother_GBps"]
        roofline = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": achieved / PEAK_HBM_GBS,
                    "traffic": measured_traffic({8: "fine_jacobi_fp64", 4: "fine_jacobi_fp32",
                                                 2: "fine_jacobi_fp16"}[word], n),
                    "kernel": f"finest-level weighted-Jacobi sweep ({wname})",
                    "bytes_per_launch": 4 * word * nnodes}
    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"bench: cpu baseline (oracle) at {args.cpu_n}^3")
        cpu = cpu_baseline(n, args.cpu_n, seeds[0], threads)
    seed_ratio = None
    if per_seed and "fp64" in mode_names:
        seed_ratio = {s: {m: v["fp64"]["ms_event_median"] / v[m]["ms_event_median"]
                          for m in v if m != "fp64"} for s, v in per_seed.items()}
    out = {
        "metric": METRIC,
        "value": hr["projections_per_s"],
        "unit": "projections/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": hr["ms_per_projection"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f64+f32",
        "data": "synthetic",
        "config": {"workload": f"variable-density periodic nodal projection {n}^3 "
                               f"(rho=1+0.5tanh(32 Fourier modes), k^-5/3 solenoidal u, "
                               f"dt=0.5h), rel. resid {args.rtol:g}",
                   "n": n, "boxes": len(case.bxs), "max_grid_size": args.max_grid_size,
                   "seed": seeds[0], "mode": head},
        "modes": results,
        "speedup_vs_fp64": (hr["projections_per_s"] / results["fp64"]["projections_per_s"]
                            if "fp64" in results else None),
        "speedup_vs_fp64_event_median": (results["fp64"]["ms_event_median"] /
                                         hr["ms_event_median"] if "fp64" in results else None),
        "seeds": per_seed,
        "seed_speedup_vs_fp64": seed_ratio,
        "check": check,
        "tuning": tuned,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    P.close()
    if net is not None:
        net.close()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
