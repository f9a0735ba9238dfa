"""The learned guess's GPU kernels against the CPU statement of the network
(oracle/nn_oracle.py, itself pinned by test_nn_oracle.py).

Bars: a single BF16 convolution or data gradient equals the float64 result
of the same BF16 operands to within half a BF16 ulp of the result plus FP32
accumulation error (|gpu - ref| <= 2^-8 |ref| + 1e-5 max|ref|); the FP32
weight gradient to 1e-4 relative L2; the whole network's output and every
layer's gradient to 1e-2 / 4e-2 relative L2 against the statement rounded
to BF16 where the device stores BF16, and to 3e-2 / 0.1 against float64.
"""
import importlib
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import nn_oracle as nn  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gs(dg):
    return importlib.import_module("3d_droplet_simulation_2007_amd.guess")


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# (grid, cinp, cin, coutp, cout, relu): the network's layer shapes, odd grids,
# a padded input and a padded output layer
CONV_CASES = [
    ((8, 8, 8), 8, 4, 16, 16, True),
    ((5, 6, 7), 16, 16, 16, 16, True),
    ((9, 4, 6), 16, 16, 32, 32, False),
    ((6, 6, 6), 32, 32, 64, 64, True),
    ((4, 4, 4), 128, 128, 64, 64, True),
    ((6, 5, 4), 96, 96, 32, 32, True),
    ((7, 7, 3), 48, 48, 16, 16, True),
    ((8, 6, 10), 16, 16, 16, 1, False),
]


# grids of at least 16 cells in x take the LDS-staged forward kernel (a
# workgroup per 32 x 2R x 2 tile with 8, 16 or 48 input channels on grids of
# 32+ cells in x, 16 x 2R x 2 otherwise): every (input groups, output tiles,
# row width) form it is built for, full tiles and tiles clipped in x, y and z
LDS_CASES = [
    ((32, 4, 2), 8, 4, 16, 16, True),
    ((35, 5, 3), 16, 16, 16, 16, True),
    ((33, 6, 5), 16, 16, 32, 32, False),
    ((34, 4, 3), 32, 32, 32, 32, True),
    ((40, 3, 3), 32, 32, 64, 64, True),
    ((38, 7, 3), 48, 48, 16, 16, True),
    ((32, 4, 3), 48, 48, 32, 32, True),
    ((33, 2, 2), 48, 40, 64, 64, True),
    ((32, 5, 2), 64, 64, 16, 16, True),
    ((36, 3, 4), 64, 64, 64, 64, True),
    ((64, 2, 3), 16, 16, 16, 1, False),
    ((33, 3, 3), 96, 96, 32, 32, True),
    ((32, 2, 3), 96, 90, 64, 64, True),
    ((34, 3, 2), 128, 128, 64, 64, True),
    ((32, 2, 2), 128, 128, 16, 16, False),
    ((16, 4, 3), 64, 64, 64, 64, True),
    ((20, 3, 2), 64, 64, 64, 64, True),
    ((17, 2, 3), 16, 16, 16, 16, True),
    ((24, 5, 3), 48, 48, 32, 32, True),
    ((16, 2, 2), 128, 128, 64, 64, True),
    ((31, 3, 3), 8, 4, 16, 16, True),
]


@pytest.mark.parametrize("case", CONV_CASES + LDS_CASES,
                         ids=[f"{c[0]}_{c[1]}to{c[3]}" for c in CONV_CASES + LDS_CASES])
def test_conv3d_forward(gs, case):
    (nx, ny, nz), cinp, cin, coutp, cout, relu = case
    print("conv fwd", case, flush=True)
    rng = np.random.default_rng(11)
    x = np.zeros((nz, ny, nx, cinp))
    x[..., :cin] = rng.normal(size=(nz, ny, nx, cin))
    w = np.zeros((coutp, 27, cinp))
    w[:cout, :, :cin] = rng.normal(scale=1.0 / np.sqrt(27 * cin), size=(cout, 27, cin))
    b = np.zeros(coutp)
    b[:cout] = rng.normal(scale=0.1, size=cout)
    xb, wb = nn.bf16(x), nn.bf16(w)
    ref = nn.conv3d(xb, wb, b, relu=relu)
    got = gs.conv3d_fwd(x, w, b, cout=cout, relu=relu).astype(np.float64)
    assert not got[..., cout:].any()
    err = np.abs(got[..., :cout] - ref[..., :cout])
    tol = 2.0 ** -8 * np.abs(ref[..., :cout]) + 1e-5 * np.abs(ref).max()
    assert (err <= tol).all(), float((err - tol).max())


@pytest.mark.parametrize("case", CONV_CASES[1:], ids=[f"{c[0]}_{c[3]}to{c[1]}" for c in CONV_CASES[1:]])
def test_conv3d_data_gradient(gs, case):
    (nx, ny, nz), cinp, cin, coutp, cout, _ = case
    print("conv dgrad", case, flush=True)
    rng = np.random.default_rng(12)
    dy = np.zeros((nz, ny, nx, coutp))
    dy[..., :cout] = rng.normal(size=(nz, ny, nx, cout))
    w = np.zeros((coutp, 27, cinp))
    w[:cout, :, :cin] = rng.normal(scale=1.0 / np.sqrt(27 * cout), size=(cout, 27, cin))
    ref = nn.conv3d_dgrad(nn.bf16(dy), nn.bf16(w))
    got = gs.conv3d_dgrad(dy, w).astype(np.float64)
    err = np.abs(got - ref)
    tol = 2.0 ** -8 * np.abs(ref) + 1e-5 * np.abs(ref).max()
    assert (err <= tol).all(), float((err - tol).max())


@pytest.mark.parametrize("case", CONV_CASES, ids=[f"{c[0]}_{c[1]}x{c[3]}" for c in CONV_CASES])
def test_conv3d_weight_gradient(gs, case):
    (nx, ny, nz), cinp, cin, coutp, cout, _ = case
    if coutp > 64:
        pytest.skip("no layer of the network has a weight gradient wider than 64")
    print("conv wgrad", case, flush=True)
    rng = np.random.default_rng(13)
    x = np.zeros((nz, ny, nx, cinp))
    x[..., :cin] = rng.normal(size=(nz, ny, nx, cin))
    dy = np.zeros((nz, ny, nx, coutp))
    dy[..., :cout] = rng.normal(size=(nz, ny, nx, cout))
    gw_ref, gb_ref = nn.conv3d_wgrad(nn.bf16(x), nn.bf16(dy))
    gw, gb = gs.conv3d_wgrad(x, dy)
    assert rel_l2(gw.astype(np.float64), gw_ref) <= 1e-4
    assert rel_l2(gb.astype(np.float64), gb_ref) <= 1e-4


def test_network_forward_and_gradient(gs):
    """Whole network at 16^3 with random weights: output and every layer's
    weight and bias gradient against (a) the statement with the device's
    BF16 storage points (bar 4e-2 relative L2: FP32 against float64
    accumulation and the odd ReLU mask or rounding that flips with it; the
    difference grows smoothly from 1e-4 at the output layer to ~1.5e-2, and
    the first layer's weight gradient, a sum of gradient x signed input with
    heavy cancellation, reaches ~3e-2) and (b) the float64 statement (bar
    0.1: what BF16 activations and gradients cost over 15 layers)."""
    print("network 16^3 forward + gradient", flush=True)
    rng = np.random.default_rng(21)
    p = nn.random_params(22).astype(np.float32)
    gs.set_params(p)
    pb = nn.bf16(p.astype(np.float64))  # the device multiplies with BF16 weights
    offs, _ = nn.param_layout()
    for (wo, bo), spec in zip(offs, nn.LAYERS):
        pb[bo:bo + spec[3]] = p[bo:bo + spec[3]]  # biases stay FP32
    params = nn.unflatten(pb)
    x0 = rng.normal(size=(16, 16, 16, 4))
    xb = nn.bf16(x0)
    target = rng.normal(size=(16, 16, 16)) * 0.1
    y = gs.forward(x0).astype(np.float64)
    y_dev, _ = nn.forward(params, xb, round_acts=True)
    y_f64, _ = nn.forward(params, xb)
    print("forward rel L2 vs device-rounded", rel_l2(y, y_dev), "vs float64", rel_l2(y, y_f64), flush=True)
    assert rel_l2(y, y_dev) <= 1e-2
    assert rel_l2(y, y_f64) <= 3e-2
    loss, g = gs.gradient(x0, target)
    loss_dev, grads_dev = nn.loss_and_grad(params, xb, target, device_rounding=True)
    loss_f64, grads_f64 = nn.loss_and_grad(params, xb, target)
    g_dev, g_f64 = nn.flatten_grads(grads_dev), nn.flatten_grads(grads_f64)
    print("loss", loss, loss_dev, loss_f64, flush=True)
    assert loss == pytest.approx(loss_dev, rel=1e-2)
    errs = []
    for i, (wo, bo) in enumerate(offs):
        cin, cinp, cout, coutp, _ = nn.LAYERS[i]
        sl_w, sl_b = slice(wo, wo + coutp * 27 * cinp), slice(bo, bo + coutp)
        row = [rel_l2(g[sl].astype(np.float64), ref[sl]) for ref in (g_dev, g_f64) for sl in (sl_w, sl_b)]
        errs.append(row)
        print(f"layer {i:2d}: vs device-rounded w {row[0]:.2e} b {row[1]:.2e};"
              f" vs float64 w {row[2]:.2e} b {row[3]:.2e}", flush=True)
    errs = np.array(errs)
    assert errs[:, :2].max() <= 4e-2, errs[:, :2].max()
    assert errs[:, 2:].max() <= 0.1, errs[:, 2:].max()


def test_network_forward_at_32cube(gs):
    """Whole-network forward at 32^3, where the finest level's five layers
    (the output layer's FP32 result among them) take the LDS-staged
    kernel: against the statement with the device's BF16 storage points,
    the same bar as at 16^3."""
    print("network 32^3 forward", flush=True)
    rng = np.random.default_rng(23)
    p = nn.random_params(24).astype(np.float32)
    gs.set_params(p)
    pb = nn.bf16(p.astype(np.float64))
    offs, _ = nn.param_layout()
    for (wo, bo), spec in zip(offs, nn.LAYERS):
        pb[bo:bo + spec[3]] = p[bo:bo + spec[3]]
    params = nn.unflatten(pb)
    x0 = rng.normal(size=(32, 32, 32, 4))
    xb = nn.bf16(x0)
    y = gs.forward(x0).astype(np.float64)
    y_dev, _ = nn.forward(params, xb, round_acts=True)
    print("forward rel L2 vs device-rounded", rel_l2(y, y_dev), flush=True)
    assert rel_l2(y, y_dev) <= 1e-2


def test_training_lowers_the_loss(gs):
    """Adam on one sample: the loss falls by 10x within 60 steps."""
    print("overfit one 16^3 sample", flush=True)
    rng = np.random.default_rng(31)
    gs.init_random(3)
    x0 = rng.normal(size=(16, 16, 16, 4)).astype(np.float32)
    target = (0.5 * x0[..., 1] + 0.2 * x0[..., 2]).astype(np.float32)
    losses = [gs.train_sample(x0, target, 2e-3) for _ in range(60)]
    print("loss", losses[0], "->", losses[-1], flush=True)
    assert losses[-1] < 0.1 * losses[0]
