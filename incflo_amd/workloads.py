"""Seeded synthetic inputs for the projection benchmarks (numpy, host side).

Variable-density single-level periodic case on [0,1]^3 (the 256^3 headline
workload):
  * rho = 1 + 0.5 tanh(f), f = sum of 32 random-phase Fourier modes with
    integer wave vectors 1 <= |k| <= 8, normalised to unit RMS;
  * u = divergence-free Gaussian random field with |u_hat| ~ |k|^(-11/6)
    (energy spectrum k^(-5/3)), normalised to max|u| = 1;
  * dt = 0.5 h / max|u|;  u* = u - dt (u.grad) u (centred differences), the
    kind of field incflo's predictor hands to the projection
    (src/incflo_apply_predictor.cpp:192-199);
  * the projection is called with scaling factor dt, so sigma = dt/rho
    (src/projection/incflo_apply_nodal_projection.cpp:117).
Arrays are (nz, ny, nx), x fastest; vectors (3, nz, ny, nx).
"""
import numpy as np


def density_field(n, seed, nmodes=32, kmax=8):
    rng = np.random.default_rng(seed)
    ks = []
    while len(ks) < nmodes:
        k = rng.integers(-kmax, kmax + 1, size=3)
        r2 = int(k @ k)
        if 1 <= r2 <= kmax * kmax:
            ks.append(k)
    amp = rng.standard_normal(nmodes)
    ph = rng.uniform(0.0, 2.0 * np.pi, nmodes)
    G = np.zeros((n, n, n), dtype=np.complex128)
    for k, a, p in zip(ks, amp, ph):
        # cell centres x = (i + 1/2)/n: cos(2 pi k.x + p) = Re exp(i(2 pi k.i/n + pi k.1/n + p))
        shift = np.pi * (k[0] + k[1] + k[2]) / n
        G[k[2] % n, k[1] % n, k[0] % n] += a * np.exp(1j * (p + shift))
    f = np.real(np.fft.ifftn(G)) * float(n) ** 3
    f /= np.sqrt(np.mean(f * f))
    return 1.0 + 0.5 * np.tanh(f)


def spectral_velocity(n, seed):
    rng = np.random.default_rng(seed + 1000003)
    k1 = np.fft.fftfreq(n, d=1.0 / n)
    kr = np.fft.rfftfreq(n, d=1.0 / n)
    KZ, KY, KX = np.meshgrid(k1, k1, kr, indexing="ij", sparse=True)
    k2 = KX * KX + KY * KY + KZ * KZ
    k2 = np.where(k2 == 0, 1.0, k2)
    amp = k2 ** (-11.0 / 12.0)
    amp[0, 0, 0] = 0.0
    shape = (n, n, n // 2 + 1)
    c = [(rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * amp for _ in range(3)]
    kd = (c[0] * KX + c[1] * KY + c[2] * KZ) / k2
    c[0] -= kd * KX
    c[1] -= kd * KY
    c[2] -= kd * KZ
    u = np.stack([np.fft.irfftn(ci, s=(n, n, n), axes=(0, 1, 2)) for ci in c])
    u /= np.max(np.sqrt(np.sum(u * u, axis=0)))
    return u


def advected_velocity(u, dt, h):
    """u* = u - dt (u.grad) u with centred periodic differences."""
    out = u.copy()
    for c in range(3):
        adv = np.zeros_like(u[c])
        for d, ax in ((0, 2), (1, 1), (2, 0)):
            dd = (np.roll(u[c], -1, axis=ax) - np.roll(u[c], 1, axis=ax)) / (2.0 * h)
            adv += u[d] * dd
        out[c] -= dt * adv
    return out


def variable_density_case(n, seed):
    """Returns (rho, u_star, dt) for the n^3 periodic variable-density case."""
    h = 1.0 / n
    rho = density_field(n, seed)
    u = spectral_velocity(n, seed)
    umax = np.max(np.sqrt(np.sum(u * u, axis=0)))  # = 1 after normalisation
    dt = 0.5 * h / max(umax, 1e-30)
    ustar = advected_velocity(u, dt, h)
    return rho, ustar, dt


def average_down(a):
    """8-cell means: (..., 2n, 2n, 2n) -> (..., n, n, n)"""
    s = a.shape
    b = a.reshape(s[:-3] + (s[-3] // 2, 2, s[-2] // 2, 2, s[-1] // 2, 2))
    return b.mean(axis=(-5, -3, -1))


def sphere_fine_boxes(nc, seed, frac=0.2, max_grid_size=32, blocking_factor=8,
                      rmin=0.05, rmax=0.15):
    """Fine boxes (refinement ratio 2, periodic unit cube of nc^3 coarse cells)
    covering seeded random spheres of radius rmin..rmax until they hold
    `frac` of the volume.  Fine cells are tagged in blocking_factor^3 blocks
    (a block whose centre lies in a sphere);
    each max_grid_size^3 chunk of the fine domain contributes the bounding box
    of its tagged blocks, so the boxes are aligned with the blocking factor and
    of uneven sizes (the kind of layout AMReX's clustering produces)."""
    rng = np.random.default_rng(seed + 7919)
    nf = 2 * nc
    bf = blocking_factor
    nb = nf // bf
    c = (np.arange(nb) + 0.5) / nb
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    inside = np.zeros((nb, nb, nb), bool)
    while inside.mean() < frac:
        p = rng.random(3)
        r = rng.uniform(rmin, rmax)
        d2 = 0.0
        for q, A in zip(p, (X, Y, Z)):
            dq = np.abs(A - q)
            dq = np.minimum(dq, 1.0 - dq)
            d2 = d2 + dq * dq
        inside |= d2 <= r * r
    tag = inside  # blocks whose centre lies in a sphere
    m = max_grid_size // bf
    boxes = []
    for K in range(0, nb, m):
        for J in range(0, nb, m):
            for I in range(0, nb, m):
                t = tag[K:K + m, J:J + m, I:I + m]
                if not t.any():
                    continue
                ks, js, is_ = (np.flatnonzero(t.any(axis=a)) for a in ((1, 2), (0, 2), (0, 1)))
                lo = (bf * (I + is_[0]), bf * (J + js[0]), bf * (K + ks[0]))
                hi = (bf * (I + is_[-1] + 1) - 1, bf * (J + js[-1] + 1) - 1, bf * (K + ks[-1] + 1) - 1)
                boxes.append((lo, hi))
    return boxes
