"""The whole incflo step on the device against the reference's own numbers
and against the oracle's step.

The decaying Taylor vortex of convergence_3d (probtype 2, [0,2]^3 periodic,
mu = 1e-3, Godunov PPM, initial projection plus 10 initial iterations, fixed
dt to t = 0.2) runs through incflo_amd.incflo.DeviceIncflo: every field stays
in device memory and every step is a C-ABI call (forces, Godunov
extrapolation, MAC projection, convective term, velocity update, implicit
viscous component solves, nodal projection).

* At 16^3 the device run equals the oracle's step (oracle/incflo_step.py) to
  1e-9 relative in u and p (both solve to 1e-11).
* At 16^3, 32^3 and 64^3 the error norms land within the bands of
  tests/test_oracle_taylor_vortex.py of convergence_3d/results_3d
  (u 8 %, p Norm0 4 %, p Norm2 1 %), and converge at second order.
"""
import json
import os
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "taylor_vortex_results_3d.json")))
MU = 1e-3


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from incflo_amd import mgx
    c = mgx.Context(0)
    yield c
    c.close()


def _device_run(ctx, n, dt):
    from incflo_amd.incflo import DeviceIncflo
    from oracle import incflo_step as S
    dx = [2.0 / n] * 3
    v0 = S.taylor_vortex_velocity((n, n, n), dx)
    I = DeviceIncflo(ctx, (n, n, n), (2.0, 2.0, 2.0), MU, v0, max_grid_size=32)
    t0 = time.perf_counter()
    I.run(0.2, dt, 10)
    wall = time.perf_counter() - t0
    vel, p = I.velocity(), I.pressure()
    e = S.errors((n, n, n), dx, vel, p, I.time, dt, MU)
    _say(f"  device n={n} steps={I.nstep} t={I.time:.4f} solves={I.solves} "
         f"wall={wall:.2f}s ({1e3 * wall / (I.nstep + 10):.2f} ms/step incl. initial iterations) "
         f"norms u {e[0]:.10g} {e[1]:.10g} p {e[2]:.10g} {e[3]:.10g}")
    I.close()
    return e, vel, p


def test_device_step_equals_oracle_step_16(ctx, orc):
    from oracle import incflo_step as S
    _say("case taylor vortex 16^3 device vs oracle")
    e, vel, p = _device_run(ctx, 16, 0.016)
    I = S.Incflo((16, 16, 16), (2.0, 2.0, 2.0), MU)
    eo = I.run(0.2, 0.016, 10)
    du = np.max(np.abs(vel - I.vel)) / np.max(np.abs(I.vel))
    dp = np.max(np.abs((p - p.mean()) - (I.p - I.p.mean()))) / np.max(np.abs(I.p))
    _say(f"  |u - u_oracle| {du:.2e}  |p - p_oracle| {dp:.2e}")
    assert du <= 1e-9 and dp <= 1e-9
    for a, b in zip(e, eo):
        assert abs(a / b - 1) <= 1e-6


@pytest.mark.parametrize("k,n,dt", [(0, 16, 0.016), (1, 32, 0.008), (2, 64, 0.004)])
def test_device_step_reproduces_reference_norms(ctx, k, n, dt):
    _say(f"case taylor vortex {n}^3 device vs convergence_3d/results_3d")
    e, _, _ = _device_run(ctx, n, dt)
    ref_u = GOLD["use_mac_phi_in_godunov_false"]["u"][k]
    ref_p = GOLD["use_mac_phi_in_godunov_false"]["p"][k]
    _say(f"  reference u {ref_u} p {ref_p}")
    assert abs(e[0] / ref_u[0] - 1) <= 0.08
    assert abs(e[1] / ref_u[1] - 1) <= 0.08
    assert abs(e[2] / ref_p[0] - 1) <= 0.04
    assert abs(e[3] / ref_p[1] - 1) <= 0.01
