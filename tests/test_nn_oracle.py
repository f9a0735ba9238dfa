"""Pins the CPU statement of the pressure-guess network (oracle/nn_oracle.py)
before it is used to check the GPU kernels: the convolution against brute
force loops, its data and weight gradients and the resampling layers by
their adjoint identities, the whole backward pass by central differences,
and the flat parameter layout by a round trip."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import nn_oracle as nn  # noqa: E402


def brute_conv(x, w, b, relu):
    nz, ny, nx, cin = x.shape
    cout = w.shape[0]
    y = np.zeros((nz, ny, nx, cout))
    for z in range(nz):
        for yy in range(ny):
            for xx in range(nx):
                acc = b.copy()
                for dz in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            zz, y2, x2 = z + dz, yy + dy, xx + dx
                            if 0 <= zz < nz and 0 <= y2 < ny and 0 <= x2 < nx:
                                t = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)
                                acc += w[:, t, :] @ x[zz, y2, x2]
                y[z, yy, xx] = np.maximum(acc, 0) if relu else acc
    return y


def test_conv_matches_brute_force():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(3, 4, 5, 3))
    w = rng.normal(size=(2, 27, 3))
    b = rng.normal(size=2)
    for relu in (False, True):
        np.testing.assert_allclose(nn.conv3d(x, w, b, relu), brute_conv(x, w, b, relu), rtol=1e-12, atol=1e-12)


def test_conv_gradients_are_adjoints():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(4, 5, 6, 3))
    w = rng.normal(size=(4, 27, 3))
    dy = rng.normal(size=(4, 5, 6, 4))
    # <conv(x), dy> = <x, dgrad(dy)>
    lhs = np.sum(nn.conv3d(x, w) * dy)
    assert np.isclose(lhs, np.sum(x * nn.conv3d_dgrad(dy, w)), rtol=1e-12)
    # <conv_w(x), dy> = <w, wgrad(x, dy)> and the bias gradient
    gw, gb = nn.conv3d_wgrad(x, dy)
    assert np.isclose(lhs, np.sum(w * gw), rtol=1e-12)
    b = rng.normal(size=4)
    assert np.isclose(np.sum(nn.conv3d(x, w, b) * dy) - lhs, np.sum(b * gb), rtol=1e-12)
    # a plausible mistake (no tap flip in the data gradient) fails the identity
    wrong = np.zeros_like(x)
    for t, (dz, ddy, ddx) in enumerate(nn.TAPS):
        wrong += nn.shifted(dy @ w[:, t, :], dz, ddy, ddx)
    assert not np.isclose(lhs, np.sum(x * wrong), rtol=1e-6)


def test_resampling_adjoints():
    rng = np.random.default_rng(3)
    f = rng.normal(size=(4, 6, 8, 5))
    c = rng.normal(size=(2, 3, 4, 5))
    assert np.isclose(np.sum(nn.pool2(f) * c), np.sum(f * nn.pool2_adj(c)), rtol=1e-12)
    assert np.isclose(np.sum(nn.up2(c) * f), np.sum(c * nn.up2_adj(f)), rtol=1e-12)
    np.testing.assert_allclose(nn.pool2(nn.up2(c)), c, rtol=1e-14)


def test_param_layout_round_trip():
    p = nn.random_params(4)
    offs, n = nn.param_layout()
    assert p.size == n
    grads = nn.unflatten(p)
    np.testing.assert_array_equal(nn.flatten_grads(grads), p)
    # padded entries are zero
    for (cin, cinp, cout, coutp, _), (wo, bo) in zip(nn.LAYERS, offs):
        blk = p[wo:wo + coutp * 27 * cinp].reshape(coutp, 27, cinp)
        assert not blk[cout:].any() and not blk[:, :, cin:].any()
        assert not p[bo + cout:bo + coutp].any()


def test_backward_matches_central_differences():
    rng = np.random.default_rng(5)
    p = nn.random_params(6)
    x0 = rng.normal(size=(8, 8, 8, 4))
    target = rng.normal(size=(8, 8, 8))
    loss, grads = nn.loss_and_grad(nn.unflatten(p), x0, target)
    g = nn.flatten_grads(grads)
    # directional derivatives along random directions restricted to a few
    # layers at a time, and along all parameters
    offs, n = nn.param_layout()
    eps = 1e-6
    groups = [range(0, 2), range(6, 8), range(12, 15), range(0, 15)]
    for layers in groups:
        d = np.zeros(n)
        for i in layers:
            wo, bo = offs[i]
            cin, cinp, cout, coutp, _ = nn.LAYERS[i]
            blk = np.zeros((coutp, 27, cinp))
            blk[:cout, :, :cin] = rng.normal(size=(cout, 27, cin))
            d[wo:wo + coutp * 27 * cinp] = blk.ravel()
            d[bo:bo + cout] = rng.normal(size=cout)
        lp = nn.loss_and_grad(nn.unflatten(p + eps * d), x0, target)[0]
        lm = nn.loss_and_grad(nn.unflatten(p - eps * d), x0, target)[0]
        fd = (lp - lm) / (2 * eps)
        assert fd == pytest.approx(float(g @ d), rel=1e-5, abs=1e-12), (list(layers), fd, float(g @ d))


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-3], dtype=np.float64)
    out = nn.bf16(vals)
    assert out[0] == 1.0
    assert out[1] == 1.0                 # tie to even (down)
    assert out[2] == 1.0 + 2 ** -7        # above the midpoint: up
    assert out[3] == -2.5
    assert abs(out[4] - 3.0e-3) <= 3.0e-3 * 2 ** -8  # half an ulp of an 8-bit significand
