"""CPU statement (numpy, float64) of the pressure-guess network of
libdroplet_gpu.so (csrc/dg_nn.hip, csrc/dg_conv.hip).

TEST INFRASTRUCTURE ONLY: tests/ and the smoke check use it as the checker
of the HIP kernels; nothing in the library or the solver path calls it.

The network has no counterpart in the reference (farzan-t/3D_Droplet_
simulation_2007 is a plain multigrid solver); this file states the same
network as the HIP code so that the GPU forward pass, data gradient, weight
gradient and the whole backward pass can be compared.  It is pinned by
tests/test_nn_oracle.py: brute-force loops for the convolution, adjoint
identities for the data and weight gradients and for the resampling
layers, and central finite differences for the whole backward pass.

Tensors are channels-last numpy arrays [nz, ny, nx, c] (x fastest in
memory, as on the device).  Convolution weights are [cout, 27, cin] with
tap t = (dz+1)*9 + (dy+1)*3 + (dx+1), zero padding at the domain edge.
"""
from __future__ import annotations

import numpy as np

# (cin, cinp, cout, coutp, relu) per layer, the library's layer table
LAYERS = [
    (4, 8, 16, 16, 1),
    (16, 16, 16, 16, 1),
    (16, 16, 32, 32, 1),
    (32, 32, 32, 32, 1),
    (32, 32, 64, 64, 1),
    (64, 64, 64, 64, 1),
    (64, 64, 64, 64, 1),
    (64, 64, 64, 64, 1),
    (128, 128, 64, 64, 1),
    (64, 64, 64, 64, 1),
    (96, 96, 32, 32, 1),
    (32, 32, 32, 32, 1),
    (48, 48, 16, 16, 1),
    (16, 16, 16, 16, 1),
    (16, 16, 1, 16, 0),
]
NFEAT = 4

TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def bf16(a):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def shifted(x, dz, dy, dx):
    """out[z,y,x] = x[z+dz, y+dy, x+dx], zero outside."""
    out = np.zeros_like(x)
    nz, ny, nx = x.shape[:3]
    zs, ys, xs = slice(max(0, -dz), nz - max(0, dz)), slice(max(0, -dy), ny - max(0, dy)), slice(max(0, -dx), nx - max(0, dx))
    zd, yd, xd = slice(max(0, dz), nz - max(0, -dz)), slice(max(0, dy), ny - max(0, -dy)), slice(max(0, dx), nx - max(0, -dx))
    out[zs, ys, xs] = x[zd, yd, xd]
    return out


def conv3d(x, w, bias=None, relu=False):
    """y[v, co] = act(bias[co] + sum_{t, ci} x[v + off(t), ci] w[co, t, ci])."""
    y = np.zeros(x.shape[:3] + (w.shape[0],))
    for t, (dz, dy, dx) in enumerate(TAPS):
        y += shifted(x, dz, dy, dx) @ w[:, t, :].T
    if bias is not None:
        y += bias
    return np.maximum(y, 0.0) if relu else y


def conv3d_dgrad(dy, w):
    """Adjoint of conv3d in x: dx[v, ci] = sum_{t, co} dy[v - off(t), co] w[co, t, ci]."""
    dx = np.zeros(dy.shape[:3] + (w.shape[2],))
    for t, (dz, ddy, ddx) in enumerate(TAPS):
        dx += shifted(dy @ w[:, t, :], -dz, -ddy, -ddx)
    return dx


def conv3d_wgrad(x, dy):
    """gw[co, t, ci] = sum_v dy[v, co] x[v + off(t), ci]; gb[co] = sum_v dy[v, co]."""
    co, ci = dy.shape[3], x.shape[3]
    gw = np.zeros((co, 27, ci))
    dyf = dy.reshape(-1, co)
    for t, (dz, ddy, ddx) in enumerate(TAPS):
        gw[:, t, :] = dyf.T @ shifted(x, dz, ddy, ddx).reshape(-1, ci)
    return gw, dyf.sum(axis=0)


def pool2(x):
    nz, ny, nx, c = x.shape
    return x.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).mean(axis=(1, 3, 5))


def pool2_adj(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 0), 2, 1), 2, 2) / 8.0


def up2(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up2_adj(g):
    nz, ny, nx, c = g.shape
    return g.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).sum(axis=(1, 3, 5))


# ---- parameters in the library's flat layout ----------------------------------

def param_layout():
    """[(woff, boff)] and the total count: per layer [coutp][27][cinp] weights
    then [coutp] biases, padded entries included (and zero)."""
    offs, off = [], 0
    for cin, cinp, cout, coutp, _ in LAYERS:
        w = off
        off += coutp * 27 * cinp
        offs.append((w, off))
        off += coutp
    return offs, off


def unflatten(p):
    """Flat vector -> list of (w [cout, 27, cin], b [cout]) without padding."""
    offs, n = param_layout()
    assert p.size == n, (p.size, n)
    out = []
    for (cin, cinp, cout, coutp, _), (wo, bo) in zip(LAYERS, offs):
        w = p[wo:wo + coutp * 27 * cinp].reshape(coutp, 27, cinp)[:cout, :, :cin]
        out.append((np.array(w, dtype=np.float64), np.array(p[bo:bo + cout], dtype=np.float64)))
    return out


def flatten_grads(grads):
    """List of (gw [cout, 27, cin], gb [cout]) -> flat vector in the library layout."""
    offs, n = param_layout()
    p = np.zeros(n)
    for (cin, cinp, cout, coutp, _), (wo, bo), (gw, gb) in zip(LAYERS, offs, grads):
        blk = np.zeros((coutp, 27, cinp))
        blk[:cout, :, :cin] = gw
        p[wo:wo + coutp * 27 * cinp] = blk.ravel()
        p[bo:bo + cout] = gb
    return p


# ---- the network ---------------------------------------------------------------

def forward(params, x0, round_acts=False):
    """x0 [nz, ny, nx, 4] -> y [nz, ny, nx] and the stored activations.
    round_acts: round every stored activation to bfloat16 as the device does
    (the products are exact in float64 either way)."""
    r = bf16 if round_acts else (lambda a: a)
    P = params
    A = {}

    def conv(i, x):
        w, b = P[i]
        return r(conv3d(x, w, b, relu=bool(LAYERS[i][4])))

    A["x0"] = x0
    A["e0a"] = conv(0, x0)
    A["e0"] = conv(1, A["e0a"])
    A["p1"] = r(pool2(A["e0"]))
    A["e1a"] = conv(2, A["p1"])
    A["e1"] = conv(3, A["e1a"])
    A["p2"] = r(pool2(A["e1"]))
    A["e2a"] = conv(4, A["p2"])
    A["e2"] = conv(5, A["e2a"])
    A["p3"] = r(pool2(A["e2"]))
    A["e3a"] = conv(6, A["p3"])
    A["e3"] = conv(7, A["e3a"])
    A["cat2"] = np.concatenate([up2(A["e3"]), A["e2"]], axis=3)
    A["d2a"] = conv(8, A["cat2"])
    A["d2"] = conv(9, A["d2a"])
    A["cat1"] = np.concatenate([up2(A["d2"]), A["e1"]], axis=3)
    A["d1a"] = conv(10, A["cat1"])
    A["d1"] = conv(11, A["d1a"])
    A["cat0"] = np.concatenate([up2(A["d1"]), A["e0"]], axis=3)
    A["d0a"] = conv(12, A["cat0"])
    A["d0"] = conv(13, A["d0a"])
    w, b = P[14]
    A["y"] = conv3d(A["d0"], w, b)[..., 0]
    return A["y"], A


def loss_and_grad(params, x0, target, device_rounding=False):
    """0.5 * mean((y - target)^2) and its gradient for every layer,
    [(gw [cout, 27, cin], gb [cout])], in float64.

    device_rounding: round to bfloat16 wherever the device stores a BF16
    tensor (activations, the output gradient, every data gradient, the
    resampling adjoints and the skip-gradient sums), so that the remaining
    difference to the device is FP32-versus-float64 accumulation."""
    r = bf16 if device_rounding else (lambda a: a)
    y, A = forward(params, x0, round_acts=device_rounding)
    V = y.size
    loss = 0.5 * float(np.mean((y - target) ** 2))
    gy = r(((y - target) / V)[..., None])
    grads = [None] * len(LAYERS)

    def back(i, x, gout, need_dx=True):
        w, b = params[i]
        grads[i] = conv3d_wgrad(x, gout)
        return r(conv3d_dgrad(gout, w)) if need_dx else None

    def relu_adj(gv, act):
        return gv * (act > 0)

    g_d0 = relu_adj(back(14, A["d0"], gy), A["d0"])
    g_d0a = relu_adj(back(13, A["d0a"], g_d0), A["d0a"])
    g_cat0 = back(12, A["cat0"], g_d0a)
    g_d1 = relu_adj(r(up2_adj(g_cat0[..., :32])), A["d1"])
    g_e0 = g_cat0[..., 32:]
    g_d1a = relu_adj(back(11, A["d1a"], g_d1), A["d1a"])
    g_cat1 = back(10, A["cat1"], g_d1a)
    g_d2 = relu_adj(r(up2_adj(g_cat1[..., :64])), A["d2"])
    g_e1 = g_cat1[..., 64:]
    g_d2a = relu_adj(back(9, A["d2a"], g_d2), A["d2a"])
    g_cat2 = back(8, A["cat2"], g_d2a)
    g_e3 = relu_adj(r(up2_adj(g_cat2[..., :64])), A["e3"])
    g_e2 = g_cat2[..., 64:]
    g_e3a = relu_adj(back(7, A["e3a"], g_e3), A["e3a"])
    g_p3 = back(6, A["p3"], g_e3a)
    g_e2 = relu_adj(r(g_e2 + pool2_adj(g_p3)), A["e2"])
    g_e2a = relu_adj(back(5, A["e2a"], g_e2), A["e2a"])
    g_p2 = back(4, A["p2"], g_e2a)
    g_e1 = relu_adj(r(g_e1 + pool2_adj(g_p2)), A["e1"])
    g_e1a = relu_adj(back(3, A["e1a"], g_e1), A["e1a"])
    g_p1 = back(2, A["p1"], g_e1a)
    g_e0 = relu_adj(r(g_e0 + pool2_adj(g_p1)), A["e0"])
    g_e0a = relu_adj(back(1, A["e0a"], g_e0), A["e0a"])
    back(0, A["x0"], g_e0a, need_dx=False)
    return loss, grads


def random_params(seed, scale=1.0):
    """He-normal weights (small output layer), zero biases, flat layout."""
    rng = np.random.default_rng(seed)
    offs, n = param_layout()
    p = np.zeros(n)
    for (cin, cinp, cout, coutp, relu), (wo, bo) in zip(LAYERS, offs):
        sd = np.sqrt(2.0 / (27 * cin)) if relu else 0.1 * np.sqrt(1.0 / (27 * cin))
        blk = np.zeros((coutp, 27, cinp))
        blk[:cout, :, :cin] = rng.normal(0.0, sd * scale, size=(cout, 27, cin))
        p[wo:wo + coutp * 27 * cinp] = blk.ravel()
        p[bo:bo + cout] = rng.normal(0.0, 0.05 * scale, size=cout)
    return p
