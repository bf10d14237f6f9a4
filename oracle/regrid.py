"""Dynamic refinement (tagging, grid checks, fine-level fill): FP64 CPU oracle.

TEST INFRASTRUCTURE ONLY (the checker for mgx_tag_cells, mgx_cluster_tags and
mgx_fill_fine_level); the product library never loads it.

tag_cells restates incflo::ErrorEst (src/incflo_tagging.cpp:60-125) on a
canonical level array with one ghost layer: rho > rhoerr; the largest
one-sided difference max(|rho_{i+1} - rho_i|, |rho_i - rho_{i-1}|) along any
direction >= gradrhoerr; the cell centre inside the closed tag region.

cons_interp restates AMReX's cell_cons_interp (CellConservativeLinear with
linear limiting; AMReX is not vendored) for refinement ratio 2, per coarse
cell and direction d:
  central slope c_q = (u_{+1} - u_{-1}) / 2, or next to an ext_dir /
    hoextrap domain face, whose ghost holds the face value,
    c_q = -16/15 u_{-1} + 1/2 u_0 + 2/3 u_{+1} - 1/10 u_{+2} (mirrored on
    the high side);
  limiter f_d = min over components q of min(1, s_q / |c_q|) with
    s_q = min(|2(u_{+1} - u_0)|, |2(u_0 - u_{-1})|), 0 where the two
    differences differ in sign (components with c_q = 0 do not limit);
  fine = u_0 + sum_d off_d f_d c_{q,d}, off_d = +-1/4.
Pinned without AMReX by what the method fixes: constants and linear fields
are reproduced exactly (also at ext_dir faces), the 8 children average to
their coarse cell, and no child leaves the range of its coarse neighbours
along each direction.  Agreement with AMReX's limiter constants beyond
these properties is parity unpinned.

check_grids tests a fine box list against what AMReX's grid generation
guarantees (every buffered tag covered, boxes disjoint, inside the domain,
aligned to the blocking factor, no side beyond max_grid_size); the cut
choices of Berger-Rigoutsos are not pinned.
"""
import numpy as np


def tag_cells(rho_g, rhoerr=None, gradrhoerr=None, region=None, prob_lo=(0, 0, 0),
              dx=(1, 1, 1)):
    """rho_g: (nz+2, ny+2, nx+2) with one ghost layer -> uint8 tags (nz, ny, nx)."""
    c = rho_g[1:-1, 1:-1, 1:-1]
    tag = np.zeros(c.shape, dtype=bool)
    if rhoerr is not None:
        tag |= c > rhoerr
    if gradrhoerr is not None:
        g = np.zeros(c.shape)
        for ax in range(3):
            sl_p = [slice(1, -1)] * 3
            sl_m = [slice(1, -1)] * 3
            sl_p[ax] = slice(2, None)
            sl_m[ax] = slice(0, -2)
            up, um = rho_g[tuple(sl_p)], rho_g[tuple(sl_m)]
            g = np.maximum(g, np.maximum(np.abs(up - c), np.abs(c - um)))
        tag |= g >= gradrhoerr
    if region is not None:
        nz, ny, nx = c.shape
        x = prob_lo[0] + (np.arange(nx) + 0.5) * dx[0]
        y = prob_lo[1] + (np.arange(ny) + 0.5) * dx[1]
        z = prob_lo[2] + (np.arange(nz) + 0.5) * dx[2]
        (xl, yl, zl), (xh, yh, zh) = region
        inx = (x >= xl) & (x <= xh)
        iny = (y >= yl) & (y <= yh)
        inz = (z >= zl) & (z <= zh)
        tag |= inz[:, None, None] & iny[None, :, None] & inx[None, None, :]
    return tag.astype(np.uint8)


def ghosted(a, periodic, ghost_lo=None, ghost_hi=None):
    """canonical (ncomp, nz, ny, nx) -> (ncomp, nz+2, ny+2, nx+2): periodic
    directions wrap; other directions take ghost planes from ghost_lo[d] /
    ghost_hi[d] (arrays of the plane's shape incl. the other ghosts) or copy
    the boundary cell when none is given."""
    out = np.pad(a, ((0, 0), (1, 1), (1, 1), (1, 1)), mode="edge")
    for d in range(3):
        ax = 3 - d
        n = a.shape[ax]
        lo = [slice(None)] * 4
        hi = [slice(None)] * 4
        lo[ax], hi[ax] = 0, n + 1
        if periodic[d]:
            src_lo = [slice(None)] * 4
            src_hi = [slice(None)] * 4
            src_lo[ax], src_hi[ax] = n, 1
            out[tuple(lo)] = out[tuple(src_lo)]
            out[tuple(hi)] = out[tuple(src_hi)]
        else:
            if ghost_lo is not None and ghost_lo[d] is not None:
                out[tuple(lo)] = ghost_lo[d]
            if ghost_hi is not None and ghost_hi[d] is not None:
                out[tuple(hi)] = ghost_hi[d]
    return out


def cons_interp(cg, periodic, bcrec=None):
    """cg: ghosted coarse (ncomp, nz+2, ny+2, nx+2) -> fine (ncomp, 2nz, 2ny, 2nx)."""
    nc = cg.shape[0]
    n = (cg.shape[3] - 2, cg.shape[2] - 2, cg.shape[1] - 2)
    bc = np.zeros((nc, 6), dtype=int) if bcrec is None else np.asarray(bcrec).reshape(nc, 6)
    u0 = cg[:, 1:-1, 1:-1, 1:-1]
    slopes = []
    for d in range(3):
        ax = 3 - d

        def sh(s):
            sl = [slice(None), slice(1, -1), slice(1, -1), slice(1, -1)]
            sl[ax] = slice(1 + s, cg.shape[ax] - 1 + s)
            return cg[tuple(sl)]

        um, up = sh(-1), sh(1)
        cen = 0.5 * (up - um)
        for q in range(nc):
            if periodic[d]:
                continue
            if bc[q, d] in (3, 4):
                idx = [q, slice(None), slice(None), slice(None)]
                idx[ax] = 0
                sl2 = [q, slice(1, -1), slice(1, -1), slice(1, -1)]
                sl2[ax] = 3  # u_{+2} of the first cell
                cen[tuple(idx)] = (-16.0 / 15.0 * um[tuple(idx)] + 0.5 * u0[tuple(idx)]
                                   + 2.0 / 3.0 * up[tuple(idx)] - 0.1 * cg[tuple(sl2)])
            if bc[q, 3 + d] in (3, 4):
                idx = [q, slice(None), slice(None), slice(None)]
                idx[ax] = n[d] - 1
                sl2 = [q, slice(1, -1), slice(1, -1), slice(1, -1)]
                sl2[ax] = n[d] - 2  # u_{-2} of the last cell
                cen[tuple(idx)] = (16.0 / 15.0 * up[tuple(idx)] - 0.5 * u0[tuple(idx)]
                                   - 2.0 / 3.0 * um[tuple(idx)] + 0.1 * cg[tuple(sl2)])
        fw, bw = 2.0 * (up - u0), 2.0 * (u0 - um)
        slp = np.where(fw * bw >= 0.0, np.minimum(np.abs(fw), np.abs(bw)), 0.0)
        with np.errstate(divide="ignore", invalid="ignore"):
            r = np.where(cen != 0.0, np.minimum(1.0, slp / np.abs(cen)), 1.0)
        f = r.min(axis=0)
        slopes.append(cen * f[None])
    out = np.empty((nc, 2 * n[2], 2 * n[1], 2 * n[0]))
    for oz in (0, 1):
        for oy in (0, 1):
            for ox in (0, 1):
                v = (u0 + (0.25 if ox else -0.25) * slopes[0]
                     + (0.25 if oy else -0.25) * slopes[1]
                     + (0.25 if oz else -0.25) * slopes[2])
                out[:, oz::2, oy::2, ox::2] = v
    return out


def buffered_tags(tags, n_error_buf, periodic):
    """TagBoxArray::buffer: every cell within n_error_buf (max norm) of a tag."""
    t = tags.astype(bool)
    out = t.copy()
    nz, ny, nx = t.shape
    zz, yy, xx = np.nonzero(t)
    b = n_error_buf
    for dk in range(-b, b + 1):
        for dj in range(-b, b + 1):
            for di in range(-b, b + 1):
                p = [xx + di, yy + dj, zz + dk]
                ok = np.ones(xx.shape, dtype=bool)
                for d, m in enumerate((nx, ny, nz)):
                    if periodic[d]:
                        p[d] = p[d] % m
                    else:
                        ok &= (p[d] >= 0) & (p[d] < m)
                out[p[2][ok], p[1][ok], p[0][ok]] = True
    return out


def check_grids(boxes, tags, n_error_buf, blocking_factor, max_grid_size, periodic):
    """Assert the fine box list (fine index space, ratio 2) covers every
    buffered coarse tag and keeps AMReX's layout rules; returns the fraction
    of fine cells that refine a buffered tag."""
    nz, ny, nx = tags.shape
    nf = (2 * nx, 2 * ny, 2 * nz)
    cov = np.zeros((2 * nz, 2 * ny, 2 * nx), dtype=np.int32)
    for lo, hi in boxes:
        for d in range(3):
            assert 0 <= lo[d] <= hi[d] < nf[d], (lo, hi)
            assert lo[d] % blocking_factor == 0 and (hi[d] + 1) % blocking_factor == 0, (lo, hi)
            assert hi[d] - lo[d] + 1 <= max_grid_size, (lo, hi)
        cov[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] += 1
    assert cov.max() <= 1, "boxes overlap"
    need = buffered_tags(tags, n_error_buf, periodic)
    fine_need = np.repeat(np.repeat(np.repeat(need, 2, 0), 2, 1), 2, 2)
    assert np.all(cov[fine_need] == 1), "a buffered tag is not covered"
    covered = cov.sum()
    return float(fine_need.sum()) / covered if covered else 1.0
