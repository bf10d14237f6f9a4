"""incflo's regrid of an AMR hierarchy (refinement ratio 2 between levels) on
device fields.  The host side is control flow; the arithmetic runs in
libmgx.so.

AmrCore::regrid as incflo drives it (src/incflo.cpp:138-197 calls regrid
every regrid_int steps):
  ErrorEst on level 0 (src/incflo_tagging.cpp:11-198)     mgx_tag_cells
  MakeNewGrids: buffer, block, cluster, chop               mgx_cluster_tags
  RemakeLevel (src/incflo_regrid.cpp:61-113) or MakeNewLevelFromCoarse
  (:9-59): velocity, density, tracer and gradp filled with
  cell_cons_interp from the old fine and the coarse level  mgx_fill_fine_level
  p_cc and p_nd of the new level set to zero (:37-38, :84-85).
No fine boxes (no tags) removes the level, as ClearLevel does.

regrid() remakes level 1 of a two-level hierarchy; regrid_levels() remakes
every level above 0 as AmrMesh::regrid(0) does (MakeNewGrids from the finest
level down, each level's tags joined by the projection of the new grids above
it so that levels stay properly nested; then the levels filled from the
bottom up).
"""
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import boxes as B
from . import mgx


@dataclass
class RegridParams:
    """incflo.rhoerr / incflo.gradrhoerr / incflo.tag_region(_lo/_hi) for
    level 0 and amr.n_error_buf / blocking_factor / max_grid_size /
    grid_eff for level 1 (None switches a tagging criterion off)."""
    rhoerr: Optional[float] = None
    gradrhoerr: Optional[float] = None
    tag_region: Optional[tuple] = None
    n_error_buf: int = 1
    blocking_factor: int = 8
    max_grid_size: int = 32
    grid_eff: float = 0.7


def fine_level_of(crse, boxes):
    """the level-1 description over the fine boxes (ratio 2, same geometry)"""
    n = [2 * crse.n_cell[d] for d in range(3)]
    return mgx.make_level(n, boxes, periodic=[crse.is_periodic[d] for d in range(3)],
                          prob_lo=[crse.prob_lo[d] for d in range(3)],
                          prob_hi=[crse.prob_hi[d] for d in range(3)])


def boxes_of(level):
    b = np.ctypeslib.as_array(level.boxes, shape=(level.nboxes * 6,))
    return [(tuple(int(v) for v in b[6 * q:6 * q + 3]), tuple(int(v) for v in b[6 * q + 3:6 * q + 6]))
            for q in range(level.nboxes)]


def regrid(ctx, crse, density, density_ng, params, crse_fields, fine_ng=None, old_fine=None,
           old_fields=None):
    """One regrid of level 1.

    crse: level-0 description; density: its boxed density with density_ng
    filled ghost cells (one at least for gradrhoerr).  crse_fields: name ->
    (boxed field, ng, ncomp, bcrec or None) on level 0; old_fields: name ->
    (boxed field, ng) on the old fine level old_fine (None: no level 1 yet).
    fine_ng: name -> ghost width of the new fine field (default: the coarse
    one).  Returns (fine level or None, name -> new boxed fine field)."""
    tags = mgx.tag_cells(ctx, crse, density, density_ng, params.rhoerr, params.gradrhoerr,
                         params.tag_region)
    boxes = mgx.cluster_tags(tags, periodic=[crse.is_periodic[d] for d in range(3)],
                             n_error_buf=params.n_error_buf,
                             blocking_factor=params.blocking_factor,
                             max_grid_size=params.max_grid_size, grid_eff=params.grid_eff)
    if not boxes:
        return None, {}
    fine = fine_level_of(crse, boxes)
    out = {}
    for name, (field, ng, ncomp, bcrec) in crse_fields.items():
        fng = ng if fine_ng is None else fine_ng.get(name, ng)
        f = ctx.array(B.field_size(boxes, fng, ncomp))
        old = None if old_fields is None else old_fields.get(name)
        mgx.fill_fine_level(ctx, crse, field, ng, ncomp, fine, f, fng, bcrec=bcrec,
                            old_fine=old_fine if old is not None else None,
                            old_field=old[0] if old is not None else None,
                            of_ng=old[1] if old is not None else 0)
        out[name] = f
    return fine, out


def project_finer(finer_boxes, n_cell, periodic, n_proper=1):
    """AmrMesh::MakeNewGrids' projection of the new level-(l+2) grids onto the
    tags of level l (n_cell: level l's extents): each box coarsened to level
    l+1, grown there by n_proper cells (periodic directions wrap, bounded ones
    stop at the domain), coarsened to level l.  Returns bool tags (nz, ny, nx)."""
    t = np.zeros((n_cell[2], n_cell[1], n_cell[0]), bool)
    for lo, hi in finer_boxes:
        idx = []
        for d in range(3):
            a = (lo[d] // 2 - n_proper) // 2
            b = (hi[d] // 2 + n_proper) // 2
            r = np.arange(a, b + 1)
            if periodic[d]:
                r = np.unique(r % n_cell[d])
            else:
                r = r[(r >= 0) & (r < n_cell[d])]
            idx.append(r)
        t[np.ix_(idx[2], idx[1], idx[0])] = True
    return t


def regrid_levels(ctx, levels, densities, params, fields, max_level, fine_ng=None,
                  fill_ghosts=None):
    """One regrid of every level above 0 (AmrMesh::regrid(0), ratio 2).

    levels: the current hierarchy, [level 0, level 1, ...] descriptions;
    densities[l] = (boxed density, ng) on level l; params: one RegridParams,
    or one per level (tagging on level l and the grids of level l+1);
    fields[l]: name -> (boxed field, ng, ncomp, bcrec or None) on level l;
    max_level: the finest level allowed.  fine_ng: name -> ghost width of the
    new fields (default: level 0's).  fill_ghosts(level, name, field, ng),
    when given, fills the ghost cells of a new field before it serves as the
    coarse field of the next level (a bounded domain needs its domain-face
    ghost cells filled there).  Returns (new levels, new fields per level);
    level 0 and its fields are returned as given."""
    L_old = len(levels) - 1
    top = min(L_old, max_level - 1)
    per = [levels[0].is_periodic[d] for d in range(3)]
    new_boxes = {}
    finer = []
    for levc in range(top, -1, -1):
        p = params[levc] if isinstance(params, (list, tuple)) else params
        rho, rng = densities[levc]
        t = mgx.tag_cells(ctx, levels[levc], rho, rng, p.rhoerr, p.gradrhoerr, p.tag_region)
        n = [levels[levc].n_cell[d] for d in range(3)]
        if finer:
            t = np.logical_or(t != 0, project_finer(finer, n, per)).astype(np.uint8)
        boxes = mgx.cluster_tags(t, periodic=per, n_error_buf=p.n_error_buf,
                                 blocking_factor=p.blocking_factor,
                                 max_grid_size=p.max_grid_size, grid_eff=p.grid_eff)
        new_boxes[levc + 1] = boxes
        finer = boxes
    new_levels, new_fields = [levels[0]], [fields[0]]
    for lev in range(1, top + 2):
        boxes = new_boxes[lev]
        if not boxes:
            break
        crse, cf = new_levels[-1], new_fields[-1]
        fine = fine_level_of(crse, boxes)
        old = levels[lev] if lev <= L_old else None
        out = {}
        for name, (field, ng, ncomp, bcrec) in cf.items():
            fng = fields[0][name][1] if fine_ng is None else fine_ng.get(name, fields[0][name][1])
            f = ctx.array(B.field_size(boxes, fng, ncomp))
            of = fields[lev].get(name) if old is not None else None
            mgx.fill_fine_level(ctx, crse, field, ng, ncomp, fine, f, fng, bcrec=bcrec,
                                old_fine=old if of is not None else None,
                                old_field=of[0] if of is not None else None,
                                of_ng=of[1] if of is not None else 0)
            if fill_ghosts is not None:
                fill_ghosts(lev, name, f, fng)
            out[name] = (f, fng, ncomp, bcrec)
        new_levels.append(fine)
        new_fields.append(out)
    return new_levels, new_fields
