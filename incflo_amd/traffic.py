"""Workload for hardware byte-counter runs of the finest-level smoother.

    rocprofv3 --pmc FETCH_SIZE -d <dir> -o run -- python3 -m incflo_amd.traffic
    rocprofv3 --pmc WRITE_SIZE -d <dir> -o run -- python3 -m incflo_amd.traffic
    python3 -m incflo_amd.traffic --summarize <dir-of-FETCH> <dir-of-WRITE>

The program first runs `--copies` launches of the plain streaming copy kernel
(mgx_stream_copy) over a buffer of known size, then one warm-up and
`--projections` nodal projections of the 256^3 variable-density case in the
given mode.  The copy launches calibrate the counters: the factor
(known bytes) / (counter reading) converts a counter reading of any kernel to
bytes.  --summarize reads rocprofv3's counter_collection CSVs and prints, per
kernel family, the average counter value per dispatch and the calibrated
bytes, in the text form kept under profiles/.
"""
import argparse
import csv
import glob
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def run(args):
    from incflo_amd import boxes as B
    from incflo_amd import mgx, workloads
    ctx = mgx.Context(0)
    nbytes = args.copy_mib << 20
    a = ctx.array(nbytes // 8)
    b = ctx.array(nbytes // 8)
    for _ in range(args.copies):
        ctx.check(mgx.lib().mgx_stream_copy(ctx.h, b.ptr, a.ptr, nbytes))
    ctx.sync()
    a.free()
    b.free()
    n = args.n
    rho, ustar, dt = workloads.variable_density_case(n, args.seed)
    bxs = B.chop_domain((n, n, n), 32)
    P = mgx.NodalProjector(ctx, mgx.make_level((n, n, n), bxs))
    vel0 = ctx.array(B.to_boxed(ustar, bxs, 1))
    vel = ctx.array(vel0.n)
    rhod = ctx.array(B.to_boxed(rho, bxs, 0))
    gp = ctx.array(B.field_size(bxs, 0, 3))
    p = ctx.array(B.field_size(bxs, 0, 1, True))
    mode = {"fp64": mgx.MODE_FP64, "mixed": mgx.MODE_MIXED}[args.mode]
    for q in range(1 + args.projections):
        vel.copy_from(vel0)
        ctx.check(mgx.lib().mgx_memset_d(ctx.h, gp.ptr, 0, gp.nbytes))
        info = P.apply_nodal_projection(vel, 1, gp, 0, p, 0, dt, density=rhod, rtol=1e-10,
                                        atol=0.0, mode=mode)
        print(f"projection {q}: {info.iters} V-cycles, rel resid {info.rel_resid:.2e}",
              flush=True)
    P.close()
    ctx.close()
    print(f"copies: {args.copies} x {nbytes} bytes read + written", flush=True)


# (label, substrings that must all occur in the demangled kernel name);
# FINE = true marks the finest level's launches of the stencil kernel
FAMILIES = [
    ("stream copy", ("stream_copy_kernel",)),
    ("fine Jacobi FP32", ("nodal_stencil_kernel<float, float, 0, true, true",)),
    ("fine Jacobi FP64", ("nodal_stencil_kernel<double, double, 0, true, true",)),
    ("update+residual FP64/FP32", ("nodal_stencil_kernel<double, float, 2,",)),
    ("update+residual FP64", ("nodal_stencil_kernel<double, double, 2,",)),
]


def _read(d):
    rows = []
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            rows += list(csv.DictReader(fh))
    return rows


def summarize(dirs, copy_bytes):
    """Per counter run: the calibration factor from the copy kernel (known
    bytes / counter bytes) and the calibrated bytes per dispatch per family."""
    out = []
    for d in dirs:
        rows = _read(d)
        if not rows:
            out.append(f"{d}: no counter_collection.csv")
            continue
        counter = rows[0]["Counter_Name"]
        per = {}
        for r in rows:
            for fam, keys in FAMILIES:
                if all(k in r["Kernel_Name"] for k in keys):
                    per.setdefault(fam, []).append(float(r["Counter_Value"]))
        copy = per.pop("stream copy", [])
        if not copy:
            out.append(f"{d}: no copy-kernel dispatches")
            continue
        kib = float(np.mean(copy))
        cal = copy_bytes / (kib * 1024.0)
        out.append(f"{d}: counter {counter}; copy kernel {kib:.1f} KiB per dispatch for "
                   f"{copy_bytes} bytes -> scale {cal:.4f}")
        for fam, vals in per.items():
            m = float(np.mean(vals))
            out.append(f"  {fam:28s} dispatches {len(vals):4d}  {counter} {m:12.1f} KiB  "
                       f"calibrated {m * 1024.0 * cal / 1e6:9.2f} MB per dispatch")
    return "\n".join(out)


def size_sweep(sizes):
    """Finest-level Jacobi sweep time against the problem size: a working set
    far beyond the 256 MB Infinity Cache (512^3 FP32: 4 arrays of 537 MB)
    shows whether the 256^3 rate is a cache effect.  The sweep is timed by the
    autotuner (2 warm-up + 5 back-to-back launches, default launch shape) on
    a constant-sigma problem; the copy kernel over the same bytes is timed
    alongside."""
    from incflo_amd import boxes as B
    from incflo_amd import mgx
    ctx = mgx.Context(0)
    for n in sizes:
        bxs = B.chop_domain((n, n, n), 64)
        for mode, word, name in ((mgx.MODE_MIXED, 4, "FP32"), (mgx.MODE_FP64, 8, "FP64")):
            P = mgx.NodalProjector(ctx, mgx.make_level((n, n, n), bxs))
            P.set_const_sigma(1.0)
            mgx.tuning_set("coarse_nodes 1\n")  # default shapes for every pass
            rep = P.autotune(mode)
            line = next(l for l in rep.splitlines()
                        if l.startswith(f"stencil {n} {n} {n} {word} {word} 0 1 "))
            f_ = line.split()
            us = float(f_[f_.index("default_us") + 1])
            nbytes = 4 * word * n ** 3
            print(f"size {n}^3 {name} fine Jacobi {us:.2f} us  {nbytes / us / 1e3:.1f} GB/s "
                  f"({nbytes / 1e6:.0f} MB algorithmic)", flush=True)
            P.close()
        # the copy kernel moving the sweep's FP32 bytes (2 arrays read+written of 2 n^3 words)
        nb = 2 * 4 * n ** 3
        a, b = ctx.array(nb // 8), ctx.array(nb // 8)
        for _ in range(2):
            ctx.check(mgx.lib().mgx_stream_copy(ctx.h, b.ptr, a.ptr, nb))
        ctx.record(0)
        for _ in range(5):
            ctx.check(mgx.lib().mgx_stream_copy(ctx.h, b.ptr, a.ptr, nb))
        ctx.record(1)
        ms = ctx.elapsed_ms(0, 1) / 5
        print(f"size {n}^3 stream copy of {nb / 1e6:.0f} MB: {1e3 * ms:.2f} us  "
              f"{2 * nb / ms / 1e6:.1f} GB/s (read + written)", flush=True)
        a.free()
        b.free()
    ctx.close()


def conv_throughput(sizes, channels):
    """The U-net's BF16-MFMA 3x3x3 convolution (mgx_conv3d_forward, FP32
    activations rounded to BF16, FP32 accumulation) on batch-1 periodic cubes:
    algorithmic 2 * 27 * cin * cout flop per voxel over the HIP-event time of
    10 back-to-back launches, against the 2.5 PFLOP/s dense BF16 peak."""
    from incflo_amd import mgx
    ctx = mgx.Context(0)
    rng = np.random.default_rng(0)
    for n in sizes:
        for c in channels:
            x = ctx.array(rng.standard_normal(n ** 3 * c).astype(np.float32), np.float32)
            w = ctx.array((0.05 * rng.standard_normal(c * 27 * c)).astype(np.float32), np.float32)
            out = ctx.array(np.zeros(n ** 3 * c, np.float32), np.float32)
            shape = (1, n, n, n)
            for _ in range(2):
                mgx.conv3d_forward(ctx, x, w, None, out, shape, c, c)
            ctx.record(0)
            for _ in range(10):
                mgx.conv3d_forward(ctx, x, w, None, out, shape, c, c)
            ctx.record(1)
            ms = ctx.elapsed_ms(0, 1) / 10
            flop = 2.0 * 27 * c * c * n ** 3
            tf = flop / ms / 1e9
            print(f"conv3d {n}^3 cin=cout={c}: {1e3 * ms:.1f} us  {tf:.1f} TFLOP/s "
                  f"({tf / 2500.0:.3f} of the 2.5 PFLOP/s BF16 dense peak)", flush=True)
            for arr in (x, w, out):
                arr.free()
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--mode", default="mixed")
    ap.add_argument("--projections", type=int, default=1)
    ap.add_argument("--copies", type=int, default=4)
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--summarize", nargs="+")
    ap.add_argument("--size-sweep", nargs="+", type=int)
    ap.add_argument("--conv", nargs="+", type=int,
                    help="time the BF16-MFMA convolution on these cube sizes")
    ap.add_argument("--conv-channels", nargs="+", type=int, default=[16, 32, 64, 128])
    args = ap.parse_args()
    if args.conv:
        conv_throughput(args.conv, args.conv_channels)
        return
    if args.size_sweep:
        size_sweep(args.size_sweep)
        return
    if args.summarize:
        print(summarize(args.summarize, args.copy_mib << 20))
    else:
        run(args)


if __name__ == "__main__":
    main()
