"""numpy helpers for the boxed field layout of include/mgx.h.

A level field is one flat array holding its boxes back to back; box b is an
AMReX Array4 over the box grown by ng ghosts (x fastest, then y, z, comp).
Canonical arrays are numpy (nz, ny, nx) for one component, (3, nz, ny, nx)
for vectors, nodal ones with nx = n (periodic) or n+1 nodes.
"""
import numpy as np


def chop_domain(n_cell, max_grid_size):
    """BoxArray(domain).maxSize(max_grid_size), in the usual lexicographic order."""
    cuts = []
    for d in range(3):
        n, m = n_cell[d], max_grid_size[d] if hasattr(max_grid_size, "__len__") else max_grid_size
        nb = -(-n // m)
        base, extra = divmod(n, nb)
        edges = [0]
        for b in range(nb):
            edges.append(edges[-1] + base + (1 if b < extra else 0))
        cuts.append(edges)
    boxes = []
    for k in range(len(cuts[2]) - 1):
        for j in range(len(cuts[1]) - 1):
            for i in range(len(cuts[0]) - 1):
                lo = (cuts[0][i], cuts[1][j], cuts[2][k])
                hi = (cuts[0][i + 1] - 1, cuts[1][j + 1] - 1, cuts[2][k + 1] - 1)
                boxes.append((lo, hi))
    return boxes


def box_dims(lo, hi, ng, nodal):
    e = 1 if nodal else 0
    return tuple(hi[d] - lo[d] + 1 + e + 2 * ng for d in range(3))


def field_size(boxes, ng, ncomp=1, nodal=False):
    return sum(ncomp * int(np.prod(box_dims(lo, hi, ng, nodal))) for lo, hi in boxes)


def to_boxed(canon, boxes, ng, nodal=False, ghost_value=None, periodic=(1, 1, 1)):
    """Scatter a canonical array (optionally with one ghost layer when
    ghost_value == 'from_canon_ghosted') into the boxed layout.

    canon: (nz,ny,nx) or (ncomp,nz,ny,nx).  For cell fields with ng>0 the ghost
    cells are filled from a ghosted canonical array if one is given via
    canon_g; here ghosts get periodic images inside the domain and 0 outside.
    """
    c = canon if canon.ndim == 4 else canon[None]
    ncomp = c.shape[0]
    out = np.zeros(field_size(boxes, ng, ncomp, nodal))
    off = 0
    nzc, nyc, nxc = c.shape[1:]
    for lo, hi in boxes:
        D = box_dims(lo, hi, ng, nodal)
        vol = int(np.prod(D))
        for comp in range(ncomp):
            blk = np.zeros((D[2], D[1], D[0]))
            idx = []
            for d, nd in zip(range(3), (nxc, nyc, nzc)):
                rng = np.arange(lo[d] - ng, lo[d] - ng + D[d])
                if periodic[d]:
                    rng = rng % nd
                    valid = np.ones(D[d], bool)
                else:
                    valid = (rng >= 0) & (rng < nd)
                    rng = np.clip(rng, 0, nd - 1)
                idx.append((rng, valid))
            (ix, vx), (iy, vy), (iz, vz) = idx
            blk = c[comp][np.ix_(iz, iy, ix)]
            mask = vz[:, None, None] & vy[None, :, None] & vx[None, None, :]
            blk = np.where(mask, blk, 0.0)
            out[off + comp * vol: off + (comp + 1) * vol] = blk.ravel()
        off += ncomp * vol
    return out


def set_ghosts_from(boxed, boxes, ng, ghosted, ncomp=3, gw=1):
    """Overwrite every cell (valid + ghost) of every box from a canonical array
    with gw ghost layers (shape (ncomp, nz+2gw, ny+2gw, nx+2gw)); ng <= gw."""
    assert ng <= gw
    off = 0
    for lo, hi in boxes:
        D = box_dims(lo, hi, ng, False)
        vol = int(np.prod(D))
        o = [lo[q] + gw - ng for q in range(3)]
        for comp in range(ncomp):
            blk = ghosted[comp, o[2]:o[2] + D[2], o[1]:o[1] + D[1], o[0]:o[0] + D[0]]
            boxed[off + comp * vol: off + (comp + 1) * vol] = blk.ravel()
        off += ncomp * vol
    return boxed


def from_boxed(boxed, boxes, ng, shape, ncomp=1, nodal=False):
    """Gather the valid region of a boxed field into a canonical array.

    shape: canonical (nz, ny, nx).  Nodal boxes contribute their nodes lo..hi+1
    (indices wrapped into the canonical array, later boxes overwrite)."""
    out = np.zeros((ncomp,) + tuple(shape))
    off = 0
    e = 1 if nodal else 0
    nz, ny, nx = shape
    for lo, hi in boxes:
        D = box_dims(lo, hi, ng, nodal)
        vol = int(np.prod(D))
        for comp in range(ncomp):
            blk = boxed[off + comp * vol: off + (comp + 1) * vol].reshape(D[2], D[1], D[0])
            v = blk[ng:D[2] - ng, ng:D[1] - ng, ng:D[0] - ng]
            iz = np.arange(lo[2], hi[2] + 1 + e) % nz
            iy = np.arange(lo[1], hi[1] + 1 + e) % ny
            ix = np.arange(lo[0], hi[0] + 1 + e) % nx
            out[comp][np.ix_(iz, iy, ix)] = v
        off += ncomp * vol
    return out if ncomp > 1 else out[0]


def face_box_dims(lo, hi, ng, d):
    return tuple(hi[e] - lo[e] + 1 + (1 if e == d else 0) + 2 * ng for e in range(3))


def face_field_size(boxes, ng, d):
    return sum(int(np.prod(face_box_dims(lo, hi, ng, d))) for lo, hi in boxes)


def faces_to_boxed(canon, boxes, ng, d):
    """canonical face array (shape (nz,ny,nx) with n_d + 1 along d) -> boxed (ghost faces 0)."""
    out = np.zeros(face_field_size(boxes, ng, d))
    off = 0
    for lo, hi in boxes:
        D = face_box_dims(lo, hi, ng, d)
        blk = np.zeros((D[2], D[1], D[0]))
        e = [1 if q == d else 0 for q in range(3)]
        blk[ng:D[2] - ng, ng:D[1] - ng, ng:D[0] - ng] = canon[lo[2]:hi[2] + 1 + e[2],
                                                              lo[1]:hi[1] + 1 + e[1],
                                                              lo[0]:hi[0] + 1 + e[0]]
        out[off:off + blk.size] = blk.ravel()
        off += blk.size
    return out


def faces_from_boxed(boxed, boxes, ng, d, shape):
    out = np.zeros(shape)
    off = 0
    for lo, hi in boxes:
        D = face_box_dims(lo, hi, ng, d)
        vol = int(np.prod(D))
        blk = boxed[off:off + vol].reshape(D[2], D[1], D[0])
        e = [1 if q == d else 0 for q in range(3)]
        out[lo[2]:hi[2] + 1 + e[2], lo[1]:hi[1] + 1 + e[1], lo[0]:hi[0] + 1 + e[0]] = \
            blk[ng:D[2] - ng, ng:D[1] - ng, ng:D[0] - ng]
        off += vol
    return out


def faces_to_boxed_n(canon, boxes, ng, d):
    """ncomp canonical face arrays (ncomp, ...) -> boxed, component-major per box."""
    nc = canon.shape[0]
    parts = []
    for lo, hi in boxes:
        per = [faces_to_boxed(canon[q], [(lo, hi)], ng, d) for q in range(nc)]
        parts.append(np.concatenate(per))
    return np.concatenate(parts)


def faces_from_boxed_n(boxed, boxes, ng, d, shape, nc):
    """inverse of faces_to_boxed_n: -> (nc,) + shape"""
    out = np.zeros((nc,) + tuple(shape))
    off = 0
    for lo, hi in boxes:
        vol = int(np.prod(face_box_dims(lo, hi, ng, d)))
        for q in range(nc):
            out[q] += 0.0
            blk = faces_from_boxed(boxed[off + q * vol: off + (q + 1) * vol], [(lo, hi)], ng, d,
                                   shape)
            e = [1 if r == d else 0 for r in range(3)]
            out[q][lo[2]:hi[2] + 1 + e[2], lo[1]:hi[1] + 1 + e[1], lo[0]:hi[0] + 1 + e[0]] = \
                blk[lo[2]:hi[2] + 1 + e[2], lo[1]:hi[1] + 1 + e[1], lo[0]:hi[0] + 1 + e[0]]
        off += nc * vol
    return out
