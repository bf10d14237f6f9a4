"""mgx_fill_boundary_noalloc, the allocation-free variant of
mgx_fill_boundary (its staging copy in a fixed device array instead of
handle memory): on two calls in a row, the second with a smaller input, it
gives exactly what mgx_fill_boundary gives; an input beyond the fixed
maximum is refused with MGX_ERR_ARG."""
import sys

import numpy as np
import pytest

from incflo_amd import boxes as B

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _junk_ghosts(u, bxs, ng):
    """boxed field with the valid cells of u and garbage in every ghost"""
    ref = B.to_boxed(u, bxs, ng)
    mask = np.zeros_like(ref)
    ncomp = u.shape[0]
    off = 0
    for lo, hi in bxs:
        D = B.box_dims(lo, hi, ng, False)
        vol = int(np.prod(D))
        for c in range(ncomp):
            blk = np.zeros((D[2], D[1], D[0]))
            blk[ng:D[2] - ng, ng:D[1] - ng, ng:D[0] - ng] = 1
            mask[off + c * vol: off + (c + 1) * vol] = blk.ravel()
        off += ncomp * vol
    return np.where(mask > 0, ref, -7.5e20)


def test_noalloc_matches_fill_boundary_twice():
    from incflo_amd import mgx
    ctx = mgx.Context(0)
    rng = np.random.default_rng(9)
    # first call: 3 components, ng 3, uneven boxes; second: 1 component on a smaller level
    for n, mgs, ng, ncomp, per in (((40, 24, 32), 16, 3, 3, (1, 1, 1)),
                                   ((12, 20, 8), 8, 2, 1, (1, 0, 1))):
        _say(f"case noalloc fill_boundary n={n} ng={ng} ncomp={ncomp} periodic={per}")
        bxs = B.chop_domain(n, mgs)
        S = mgx.Step(ctx, mgx.make_level(n, bxs, periodic=per))
        u = rng.standard_normal((ncomp, n[2], n[1], n[0]))
        junk = _junk_ghosts(u, bxs, ng)
        a = ctx.array(junk)
        b = ctx.array(junk)
        S.fill_boundary(a, ng, ncomp)
        S.fill_boundary_noalloc(b, ng, ncomp)
        ra, rb = a.download(), b.download()
        assert np.array_equal(ra, rb)
        if per == (1, 1, 1):
            assert np.array_equal(rb, B.to_boxed(u, bxs, ng))
        S.close()
    ctx.close()


def test_noalloc_refuses_beyond_the_fixed_maximum():
    from incflo_amd import mgx
    ctx = mgx.Context(0)
    n = (144, 144, 144)  # 3 x 144^3 = 8,957,952 > 8,388,608 doubles
    bxs = B.chop_domain(n, 72)
    S = mgx.Step(ctx, mgx.make_level(n, bxs))
    f = ctx.array(B.field_size(bxs, 1, 3))
    with pytest.raises(mgx.MgxError) as e:
        S.fill_boundary_noalloc(f, 1, 3)
    assert e.value.code == mgx.MGX_ERR_ARG
    S.fill_boundary_noalloc(f, 1, 2)  # 2 components fit
    S.close()
    ctx.close()
