"""Pins the oracle's 3x3x3 convolution and U-net (oracle/orc_unet.c).

The learned predictor has no counterpart in the reference, so these checks
rest on independent implementations and exact identities:
  * forward convolution == scipy.ndimage.correlate (wrap / zero padding);
  * data and weight gradients == adjoints of the (bi)linear forward map;
  * bfloat16 rounding == round-to-nearest-even known answers;
  * whole-network gradients == central finite differences of the forward.
"""
import numpy as np
import pytest
import scipy.ndimage as ndi

import oracle


def _conv_scipy(x, w, bias, cout, periodic):
    batch, nz, ny, nx, cin = x.shape
    k = w.reshape(cout, 3, 3, 3, cin)  # (co, dz, dy, dx, ci)
    out = np.zeros((batch, nz, ny, nx, cout))
    mode = "wrap" if periodic else "constant"
    for b in range(batch):
        for co in range(cout):
            acc = np.full((nz, ny, nx), bias[co] if bias is not None else 0.0)
            for ci in range(cin):
                acc += ndi.correlate(x[b, ..., ci], k[co, ..., ci], mode=mode, cval=0.0)
            out[b, ..., co] = acc
    return out


@pytest.mark.parametrize("periodic", [True, False])
def test_conv_forward_matches_scipy(orc, periodic):
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 4, 5, 6, 3))
    w = rng.standard_normal((2, 27 * 3))
    b = rng.standard_normal(2)
    per = (1, 1, 1) if periodic else (0, 0, 0)
    got = oracle.conv3d_forward(x, w, b, 2, periodic=per)
    ref = _conv_scipy(x, w, b, 2, periodic)
    assert np.allclose(got, ref, rtol=1e-13, atol=1e-12)
    relu = oracle.conv3d_forward(x, w, b, 2, periodic=per, relu=True)
    assert np.array_equal(relu, np.maximum(got, 0.0))


def test_conv_mixed_periodicity_taps(orc):
    # tap order: tap = (dz+1)*9 + (dy+1)*3 + (dx+1); a single +x tap shifts by one voxel
    x = np.arange(4 * 4 * 4, dtype=np.float64).reshape(1, 4, 4, 4, 1)
    w = np.zeros((1, 27))
    w[0, 1 * 9 + 1 * 3 + 2] = 1.0  # dx = +1
    out = oracle.conv3d_forward(x, w, None, 1, periodic=(1, 0, 0))
    assert np.array_equal(out[0, ..., 0], np.roll(x[0, ..., 0], -1, axis=2))
    out = oracle.conv3d_forward(x, w, None, 1, periodic=(0, 1, 1))
    ref = np.roll(x[0, ..., 0], -1, axis=2)
    ref[:, :, -1] = 0.0
    assert np.array_equal(out[0, ..., 0], ref)


@pytest.mark.parametrize("periodic", [(1, 1, 1), (0, 1, 0)])
def test_conv_backward_is_adjoint(orc, periodic):
    rng = np.random.default_rng(5)
    cin, cout = 4, 3
    x = rng.standard_normal((2, 4, 3, 5, cin))
    w = rng.standard_normal((cout, 27 * cin))
    d = rng.standard_normal((2, 4, 3, 5, cout))
    y = oracle.conv3d_forward(x, w, None, cout, periodic=periodic)
    din, dw, db = oracle.conv3d_backward(x, w, d, periodic=periodic)
    lhs = np.sum(y * d)
    # linear in x: <conv(x), d> = <x, conv^T d>; linear in w: <conv(x;w), d> = <w, dw>
    assert abs(lhs - np.sum(x * din)) < 1e-11 * max(1.0, abs(lhs))
    assert abs(lhs - np.sum(w * dw)) < 1e-11 * max(1.0, abs(lhs))
    assert np.allclose(db, d.reshape(-1, cout).sum(axis=0), rtol=1e-13, atol=1e-12)


def test_bf16_rounding_known_answers(orc):
    # identity kernel (centre tap) returns the rounded input
    vals = np.array([1.0 + 2.0 ** -8,              # tie -> even (1.0)
                     1.0 + 3 * 2.0 ** -8,          # tie -> even (1 + 2^-6)
                     1.0 + 2.0 ** -8 + 2.0 ** -20,  # above tie -> 1 + 2^-7
                     -3.0 - 2.0 ** -7 - 2.0 ** -12,
                     0.1])
    x = np.zeros((1, 1, 1, len(vals), 1))
    x[0, 0, 0, :, 0] = vals
    w = np.zeros((1, 27))
    w[0, 13] = 1.0
    out = oracle.conv3d_forward(x, w, None, 1, periodic=(0, 0, 0), bf16=True)[0, 0, 0, :, 0]
    exp = [1.0, 1.0 + 2.0 ** -6, 1.0 + 2.0 ** -7, -3.0 - 2.0 ** -6, 0.10009765625]
    assert out.tolist() == exp
    # values already representable in bfloat16 are unchanged
    rng = np.random.default_rng(1)
    xr = rng.integers(-64, 64, size=(1, 3, 3, 3, 2)) / 16.0
    wr = rng.integers(-64, 64, size=(2, 54)) / 32.0
    assert np.array_equal(oracle.conv3d_forward(xr, wr, None, 2, bf16=True),
                          oracle.conv3d_forward(xr, wr, None, 2, bf16=False))


def test_unet_gradient_finite_differences(orc):
    rng = np.random.default_rng(11)
    base, levels = 16, 4
    n = oracle.unet_nparam(base, levels, 8)
    # parameter count of the documented architecture
    convs = [(8, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (64, 128), (128, 128),
             (192, 64), (64, 64), (96, 32), (32, 32), (48, 16), (16, 16)]
    assert n == sum(co * 27 * ci + co for ci, co in convs) + base + 1
    p = np.zeros(n)
    off = 0
    for ci, co in convs:
        p[off:off + co * 27 * ci] = rng.standard_normal(co * 27 * ci) * np.sqrt(2.0 / (27 * ci))
        off += co * 27 * ci
        p[off:off + co] = 0.05 * rng.standard_normal(co)
        off += co
    p[off:off + base] = rng.standard_normal(base) / 4
    p[-1] = 0.1
    x = np.zeros((1, 8, 8, 8, 8))
    x[..., :2] = rng.standard_normal((1, 8, 8, 8, 2))
    d = rng.standard_normal((1, 8, 8, 8))
    g = oracle.unet_grad(p, x, d, base, levels)
    # probe parameters in every layer, including biases and the head
    idx = []
    off = 0
    for ci, co in convs:
        idx.append(off + int(rng.integers(co * 27 * ci)))
        off += co * 27 * ci
        idx.append(off + int(rng.integers(co)))
        off += co
    idx += [off + 3, n - 1]
    h = 1e-6
    for i in idx:
        pp = p.copy(); pp[i] += h
        pm = p.copy(); pm[i] -= h
        fd = (np.sum(d * oracle.unet_forward(pp, x, base, levels)) -
              np.sum(d * oracle.unet_forward(pm, x, base, levels))) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])
