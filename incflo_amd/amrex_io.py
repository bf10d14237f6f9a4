"""Readers for the AMReX formats libmgx.so writes (plotfiles, VisMF
multifabs, incflo checkpoint headers), for checking and post-processing
(numpy; the same layouts amrex::VisMF and amrex::WriteMultiLevelPlotfile
produce: see include/mgx.h's output section).
"""
import os
import re

import numpy as np

_BOX = re.compile(r"\(\((-?\d+),(-?\d+),(-?\d+)\) \((-?\d+),(-?\d+),(-?\d+)\) \((\d+),(\d+),(\d+)\)\)")


def parse_boxes(text):
    """all "((lo) (hi) (type))" boxes in text, as (lo, hi, type) tuples"""
    out = []
    for m in _BOX.finditer(text):
        v = [int(x) for x in m.groups()]
        out.append((tuple(v[0:3]), tuple(v[3:6]), v[6]))
    return out


def read_multifab(prefix):
    """VisMF multifab -> dict(ncomp, ngrow, boxes, fabs): fabs[b] is an array
    (ncomp, nz, ny, nx) over box b grown by ngrow; also the header's per-fab
    minima / maxima."""
    with open(prefix + "_H") as f:
        lines = f.read().split("\n")
    vers, how, ncomp, ngrow = (int(lines[q]) for q in range(4))
    nb = int(lines[4].lstrip("(").split()[0])  # BoxArray: "(n 0", n boxes, ")"
    boxes = parse_boxes("\n".join(lines[5:5 + nb]))
    assert len(boxes) == nb and lines[5 + nb] == ")"
    rest = lines[6 + nb:]
    nfod = int(rest[0])
    fods = []
    for q in range(nfod):
        tag, name, off = rest[1 + q].split()
        assert tag == "FabOnDisk:"
        fods.append((name, int(off)))
    vals = [l for l in rest[1 + nfod:] if l.strip()]
    mins = [[float(x) for x in vals[1 + b].rstrip(",").split(",")] for b in range(nfod)]
    maxs = [[float(x) for x in vals[2 + nfod + b].rstrip(",").split(",")] for b in range(nfod)]
    d = os.path.dirname(prefix)
    fabs = []
    for (lo, hi, t), (name, off) in zip(boxes, fods):
        with open(os.path.join(d, name), "rb") as f:
            f.seek(off)
            head = f.readline().decode()
            assert head.startswith("FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))"), head
            fb = parse_boxes(head)[0]
            nc = int(head.split()[-1])
            dims = [fb[1][q] - fb[0][q] + 1 for q in range(3)]
            a = np.fromfile(f, dtype="<f8", count=nc * int(np.prod(dims)))
        fabs.append(a.reshape(nc, dims[2], dims[1], dims[0]))
    return {"version": vers, "how": how, "ncomp": ncomp, "ngrow": ngrow, "boxes": boxes,
            "fabs": fabs, "min": mins, "max": maxs}


def to_canonical(mf, shape, nodal=False, ncomp=None):
    """valid regions of a read multifab -> canonical (ncomp, nz, ny, nx)"""
    nc = mf["ncomp"] if ncomp is None else ncomp
    g = mf["ngrow"]
    out = np.zeros((nc,) + tuple(shape))
    nz, ny, nx = shape
    for (lo, hi, t), a in zip(mf["boxes"], mf["fabs"]):
        v = a[:nc, g:a.shape[1] - g, g:a.shape[2] - g, g:a.shape[3] - g]
        iz = np.arange(lo[2], hi[2] + 1) % nz
        iy = np.arange(lo[1], hi[1] + 1) % ny
        ix = np.arange(lo[0], hi[0] + 1) % nx
        out[:, iz[:, None, None], iy[None, :, None], ix[None, None, :]] = v
    return out


def read_plotfile(path):
    """HyperCLaw-V1.1 plotfile -> dict(names, time, finest_level, prob_lo,
    prob_hi, domains, dx, levels) with levels[l] the level's multifab"""
    with open(os.path.join(path, "Header")) as f:
        L = [l.rstrip("\n") for l in f]
    assert L[0] == "HyperCLaw-V1.1", L[0]
    nv = int(L[1])
    names = L[2:2 + nv]
    q = 2 + nv
    dim = int(L[q]); time = float(L[q + 1]); finest = int(L[q + 2])
    prob_lo = [float(x) for x in L[q + 3].split()]
    prob_hi = [float(x) for x in L[q + 4].split()]
    domains = parse_boxes(L[q + 6])
    dx = [[float(x) for x in L[q + 8 + l].split()] for l in range(finest + 1)]
    levels = []
    for l in range(finest + 1):
        levels.append(read_multifab(os.path.join(path, f"Level_{l}", "Cell")))
    return {"names": names, "dim": dim, "time": time, "finest_level": finest,
            "prob_lo": prob_lo, "prob_hi": prob_hi, "domains": domains, "dx": dx,
            "levels": levels}
