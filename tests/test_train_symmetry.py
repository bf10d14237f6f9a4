"""The training-set augmentation maps solutions of the periodic nodal problem
to solutions: for every cube symmetry applied to (rhs at nodes, log sigma at
cells, phi at nodes), the oracle's operator still gives L_sigma' phi' = b'."""
import numpy as np
import pytest

from incflo_amd.train_unet import PERMS, symmetry


@pytest.mark.parametrize("perm", PERMS)
@pytest.mark.parametrize("flips", [(0, 0, 0), (1, 0, 0), (0, 1, 1), (1, 1, 1)])
def test_symmetry_preserves_the_equation(orc, perm, flips):
    n = 8
    g = orc.Geom.make((n, n, n))
    rng = np.random.default_rng(7)
    sigma = np.exp(0.3 * rng.standard_normal(g.cell_shape()))
    phi = rng.standard_normal(g.node_shape())
    b = orc.nodal_apply(g, sigma, phi)
    x = np.stack([b, np.log(sigma)], axis=-1).astype(np.float32)
    y, t = symmetry(x, phi.astype(np.float32), perm, flips)
    b2 = y[..., 0].astype(np.float64)
    s2 = np.exp(y[..., 1].astype(np.float64))
    r = orc.nodal_apply(g, s2, t.astype(np.float64)) - b2
    assert np.max(np.abs(r)) <= 1e-5 * np.max(np.abs(b))
    # a reflection that moved cells like nodes would break the equation
    if any(flips):
        bad = np.exp(np.roll(y[..., 1], 1, axis=int(np.argmax(flips))).astype(np.float64))
        rb = orc.nodal_apply(g, bad, t.astype(np.float64)) - b2
        assert np.max(np.abs(rb)) > 1e-3 * np.max(np.abs(b))
