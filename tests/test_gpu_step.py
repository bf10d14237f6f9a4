"""GPU parity of the predictor-step entry points (mgx_step_*) against numpy
restatements of the reference's elementwise updates, on uneven boxes:
FillBoundary with periodic images, compute_vel_forces
(src/incflo_compute_forces.cpp:134-148), update_velocity predictor in the
velocity and momentum forms (src/incflo_update_velocity.cpp:14-143) and
update_density (src/incflo_update_density.cpp:14-64).  These are elementwise
FP64 formulas: the bar is 1e-14 relative."""
import sys

import numpy as np
import pytest

from incflo_amd import boxes as B

pytestmark = pytest.mark.gpu

N = (20, 12, 16)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def env():
    from incflo_amd import mgx
    c = mgx.Context(0)
    bxs = B.chop_domain(N, 8)
    S = mgx.Step(c, mgx.make_level(N, bxs))
    yield c, bxs, S
    S.close()
    c.close()


def _cells(rng, ncomp=3):
    return rng.standard_normal((ncomp, N[2], N[1], N[0]))


def test_fill_boundary_periodic(env):
    ctx, bxs, S = env
    _say("case fill_boundary ng=3")
    rng = np.random.default_rng(1)
    u = _cells(rng)
    ref = B.to_boxed(u, bxs, 3)  # ghosts = periodic images
    # destroy every ghost: keep only the valid cells
    valid = ref.copy()
    mask = np.zeros_like(ref)
    off = 0
    for lo, hi in bxs:
        D = B.box_dims(lo, hi, 3, False)
        vol = int(np.prod(D))
        for c in range(3):
            blk = np.zeros((D[2], D[1], D[0]))
            blk[3:-3, 3:-3, 3:-3] = 1
            mask[off + c * vol: off + (c + 1) * vol] = blk.ravel()
        off += 3 * vol
    junk = np.where(mask > 0, valid, 1e30)
    d = ctx.array(junk)
    S.fill_boundary(d, 3, 3)
    assert np.array_equal(d.download(), ref)


def test_forces_and_velocity_update(env):
    ctx, bxs, S = env
    _say("case forces / update_velocity / update_density")
    rng = np.random.default_rng(2)
    gp, div, uo, conv = _cells(rng), _cells(rng), _cells(rng), _cells(rng)
    rho = 1.0 + 0.5 * rng.random((N[2], N[1], N[0]))
    g, gp0 = (0.1, -0.2, -9.8), (0.3, 0.0, -0.1)
    f = ctx.array(B.field_size(bxs, 1, 3))
    gpd = ctx.array(B.to_boxed(gp, bxs, 0))
    rd = ctx.array(B.to_boxed(rho, bxs, 2))
    dd = ctx.array(B.to_boxed(div, bxs, 0))
    S.compute_vel_forces(gpd, 0, rd, 2, f, 1, divtau=dd, divtau_ng=0, gravity=g, gp0=gp0)
    got = B.from_boxed(f.download(), bxs, 1, rho.shape, ncomp=3)
    ref = np.stack([-(gp[c] + gp0[c]) / rho + g[c] + div[c] for c in range(3)])
    assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref))
    S.compute_vel_forces(None, 0, rd, 2, f, 1, gravity=g, gp0=gp0,
                         include_pressure_gradient=False)
    got = B.from_boxed(f.download(), bxs, 1, rho.shape, ncomp=3)
    ref = np.stack([-gp0[c] / rho + g[c] for c in range(3)])
    assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref))
    # velocity form with the explicit divtau coefficient
    fo = _cells(rng)
    fd = ctx.array(B.to_boxed(fo, bxs, 1))
    uod = ctx.array(B.to_boxed(uo, bxs, 3))
    cd = ctx.array(B.to_boxed(conv, bxs, 0))
    vd = ctx.array(B.field_size(bxs, 3, 3))
    dt = 0.013
    S.update_velocity(vd, 3, uod, 3, cd, 0, fd, 1, dt, divtau=dd, divtau_ng=0, divtau_coef=0.5)
    got = B.from_boxed(vd.download(), bxs, 3, rho.shape, ncomp=3)
    ref = uo + dt * (conv + fo + 0.5 * div)
    assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref))
    # momentum form
    ro, rn = 1.0 + rng.random(rho.shape), 1.0 + rng.random(rho.shape)
    rh = 0.5 * (ro + rn)
    rod, rnd, rhd = (ctx.array(B.to_boxed(a, bxs, 1)) for a in (ro, rn, rh))
    S.update_velocity(vd, 3, uod, 3, cd, 0, fd, 1, dt, rho_old=rod, rho_new=rnd, rho_nph=rhd,
                      rho_ng=1)
    got = B.from_boxed(vd.download(), bxs, 3, rho.shape, ncomp=3)
    ref = (ro * uo + dt * (conv + rh * fo)) / rn
    assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref))
    # density
    drdt = rng.standard_normal(rho.shape)
    drd = ctx.array(B.to_boxed(drdt, bxs, 0))
    rnew = ctx.array(B.field_size(bxs, 1, 1))
    rnph = ctx.array(B.field_size(bxs, 0, 1))
    S.update_density(rod, 1, drd, 0, dt, rnew, 1, rnph, 0)
    rn2 = B.from_boxed(rnew.download(), bxs, 1, rho.shape)
    rh2 = B.from_boxed(rnph.download(), bxs, 0, rho.shape)
    assert np.max(np.abs(rn2 - (ro + dt * drdt))) <= 1e-15 * np.max(np.abs(ro))
    assert np.max(np.abs(rh2 - 0.5 * (ro + ro + dt * drdt))) <= 1e-15 * np.max(np.abs(ro))
