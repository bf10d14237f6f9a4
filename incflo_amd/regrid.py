"""incflo's regrid of a two-level hierarchy (refinement ratio 2) on device
fields.  The host side is control flow; the arithmetic runs in libmgx.so.

AmrCore::regrid as incflo drives it (src/incflo.cpp:138-197 calls regrid
every regrid_int steps):
  ErrorEst on level 0 (src/incflo_tagging.cpp:11-198)     mgx_tag_cells
  MakeNewGrids: buffer, block, cluster, chop               mgx_cluster_tags
  RemakeLevel (src/incflo_regrid.cpp:61-113) or MakeNewLevelFromCoarse
  (:9-59): velocity, density, tracer and gradp filled with
  cell_cons_interp from the old fine and the coarse level  mgx_fill_fine_level
  p_cc and p_nd of the new level set to zero (:37-38, :84-85).
No fine boxes (no tags) removes the level, as ClearLevel does.
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
