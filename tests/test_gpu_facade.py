"""The C++ facade (include/mgx_amrex.H) driven the way incflo drives
AMReX-Hydro's projectors (tests/cpp/facade_check.cpp) gives bit-identical
results to the same calls made directly through the C ABI, and its nodal
projection meets the oracle's 1e-10 FP64 parity."""
import os
import subprocess
import sys

import numpy as np
import pytest

from incflo_amd import boxes as B

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "incflo_amd", "_build", "facade_check")


@pytest.fixture(scope="module")
def ctx():
    from incflo_amd import mgx
    c = mgx.Context(0)
    yield c
    c.close()


def _periodic_faces(rng, n):
    u = []
    for d in range(3):
        shape = [n, n, n]
        shape[2 - d] += 1
        a = rng.standard_normal(shape)
        sl_last = [slice(None)] * 3
        sl_first = [slice(None)] * 3
        sl_last[2 - d] = n
        sl_first[2 - d] = 0
        a[tuple(sl_last)] = a[tuple(sl_first)]
        u.append(a)
    return u


def test_facade_matches_direct_abi_and_oracle(ctx, orc, tmp_path):
    from incflo_amd import mgx, workloads
    n, mgs = 32, 16
    print(f"case facade n={n}", flush=True)
    assert os.path.exists(DRIVER), "build the library first (incflo_amd.build)"
    rho, ustar, dt = workloads.variable_density_case(n, 3)
    sigma = dt / rho
    bxs = B.chop_domain((n, n, n), mgs)
    vel_b = B.to_boxed(ustar, bxs, 1)
    sig_b = B.to_boxed(sigma, bxs, 0)
    rng = np.random.default_rng(5)
    umac = _periodic_faces(rng, n)
    umac_b = [B.faces_to_boxed(umac[d], bxs, 0, d) for d in range(3)]
    vel_b.tofile(tmp_path / "vel.bin")
    sig_b.tofile(tmp_path / "sigma.bin")
    for d in range(3):
        umac_b[d].tofile(tmp_path / f"umac{d}.bin")
    r = subprocess.run([DRIVER, str(tmp_path), str(n), str(mgs)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "facade_check ok" in r.stdout

    # the same calls through the C ABI
    lev = mgx.make_level((n, n, n), bxs)
    P = mgx.NodalProjector(ctx, lev)
    P.set_sigma(ctx.array(sig_b), 0)
    vd = ctx.array(vel_b)
    info = P.project(vd, 1, 1e-12, 0.0)
    phid = ctx.array(B.field_size(bxs, 0, 1, True))
    gpd = ctx.array(B.field_size(bxs, 0, 3))
    P.getPhi(phid, 0)
    P.getGradPhi(gpd, 0)
    assert np.array_equal(np.fromfile(tmp_path / "vel_out.bin"), vd.download())
    assert np.array_equal(np.fromfile(tmp_path / "phi.bin"), phid.download())
    assert np.array_equal(np.fromfile(tmp_path / "gphi.bin"), gpd.download())
    P.close()
    M = mgx.MacProjector(ctx, lev)
    M.updateBeta(0.25)
    ud = [ctx.array(umac_b[d]) for d in range(3)]
    mp = ctx.array(B.field_size(bxs, 0, 1))
    M.project(ud[0], ud[1], ud[2], 0, 1e-10, 0.0, phi=mp, phi_ng=0)
    for d in range(3):
        assert np.array_equal(np.fromfile(tmp_path / f"umac_out{d}.bin"), ud[d].download())
    assert np.array_equal(np.fromfile(tmp_path / "macphi.bin"), mp.download())
    M.close()

    # FP64 parity of the facade's nodal projection with the oracle
    g = orc.Geom.make((n, n, n))
    ug = np.zeros(g.ghost_vec_shape())
    ug[:, 1:-1, 1:-1, 1:-1] = ustar
    orc.fill_vel_ghosts(g, ug)
    phio, gphio, io, ret = orc.nodal_project(g, sigma, ug, rtol=1e-12)
    assert ret == 0
    phi = B.from_boxed(np.fromfile(tmp_path / "phi.bin"), bxs, 0, g.node_shape(), nodal=True)
    phi -= phi.mean()
    phio = phio - phio.mean()
    assert np.max(np.abs(phi - phio)) <= 1e-10 * np.max(np.abs(phio))
    vel = B.from_boxed(np.fromfile(tmp_path / "vel_out.bin"), bxs, 1, g.cell_shape(), ncomp=3)
    uo = ug[:, 1:-1, 1:-1, 1:-1]
    assert np.max(np.abs(vel - uo)) <= 1e-10 * np.max(np.abs(uo))
