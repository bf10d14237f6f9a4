"""The oracle's whole incflo step against the reference's own numbers.

convergence_3d/results_3d (committed as tests/golden/taylor_vortex_results_3d.json)
holds the error norms incflo printed for the decaying Taylor vortex
(probtype 2, [0,2]^3 periodic, mu = 1e-3, Godunov PPM,
use_mac_phi_in_godunov = false, initial projection plus 10 initial
iterations, fixed dt to t = 0.2).  oracle/incflo_step.py composes the same
sequence from the oracle's pieces; its norms must land within a stated band
of the reference's:

  u: 8 % (Norm0 and Norm2)      p: 4 % (Norm0), 1 % (Norm2)

The bands cover what this build does not pin against AMReX / AMReX-Hydro,
which are not vendored: the PPM limiter's branch details, the viscous
operator (the reference default is the tensor solve, this build solves the
components separately) and MLMG's own iteration path.  Measured: 16^3
u 0.8 % / 2.9 %, p 0.6 % / 0.15 %; 32^3 u 6.9 % / 3.5 %, p 2.6 % / 0.17 %;
64^3 (8 minutes on the CPU, so not run here; the device step runs it)
u 1.5 % / 2.9 %, p 1.5 % / 0.3 %.
The second-order convergence of the scheme is checked as well.
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "taylor_vortex_results_3d.json")))
BAND_U, BAND_P0, BAND_P2 = 0.08, 0.04, 0.01


def _run(orc, n, dt):
    from oracle import incflo_step as S
    I = S.Incflo((n, n, n), (2.0, 2.0, 2.0), 1e-3)
    return I.run(0.2, dt, 10)


def _check(got, k):
    ref_u = GOLD["use_mac_phi_in_godunov_false"]["u"][k]
    ref_p = GOLD["use_mac_phi_in_godunov_false"]["p"][k]
    u0, u2, p0, p2 = got
    assert abs(u0 / ref_u[0] - 1) <= BAND_U, (u0, ref_u)
    assert abs(u2 / ref_u[1] - 1) <= BAND_U, (u2, ref_u)
    assert abs(p0 / ref_p[0] - 1) <= BAND_P0, (p0, ref_p)
    assert abs(p2 / ref_p[1] - 1) <= BAND_P2, (p2, ref_p)


def test_taylor_vortex_16(orc):
    got = _run(orc, 16, 0.016)
    print("16^3 oracle norms", got)
    _check(got, 0)


@pytest.mark.slow
def test_taylor_vortex_32_and_order(orc):
    g16 = _run(orc, 16, 0.016)
    g32 = _run(orc, 32, 0.008)
    print("32^3 oracle norms", g32)
    _check(g32, 1)
    # second order in the L2 norms, like the reference's own numbers
    assert np.log2(g16[1] / g32[1]) > 1.8
    assert np.log2(g16[3] / g32[3]) > 1.8
