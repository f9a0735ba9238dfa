"""Python host side of the learned initial guess (dg_guess_* in
include/droplet_gpu.h): network parameters, single-layer and whole-network
calls on host arrays, the device-resident sample store and training.

The network, its training and its use inside the pressure solve all run in
libdroplet_gpu.so; this module only moves arrays and files.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .api import DropletError, _check, _i, lib

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "weights", "guess_unet.npz")

_fp = ctypes.POINTER(ctypes.c_float)


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != tuple(shape):
        raise DropletError(f"expected shape {tuple(shape)}, got {a.shape}")
    return a


def _pf(a):
    return a.ctypes.data_as(_fp)


def nparams():
    n = ctypes.c_int(0)
    _check(lib().dg_guess_nparams(ctypes.byref(n)), "dg_guess_nparams")
    return n.value


def init_random(seed=0):
    _check(lib().dg_guess_init_random(_i(seed)), "dg_guess_init_random")


def set_params(p):
    p = _f32(p).ravel()
    _check(lib().dg_guess_set_params(_pf(p), _i(p.size)), "dg_guess_set_params")


def get_params():
    n = nparams()
    p = np.empty(n, dtype=np.float32)
    _check(lib().dg_guess_get_params(_pf(p), _i(n)), "dg_guess_get_params")
    return p


def save_weights(path, meta=None):
    """Current parameters to an .npz (array 'params' plus string metadata)."""
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    extra = {k: np.array(str(v)) for k, v in (meta or {}).items()}
    np.savez(path, params=get_params(), **extra)


def load_weights(path=DEFAULT_WEIGHTS):
    with np.load(path, allow_pickle=False) as z:
        set_params(z["params"])
        return {k: str(z[k]) for k in z.files if k != "params"}


def forward(feat):
    """feat [nz, ny, nx, 4] -> prediction [nz, ny, nx] (FP32)."""
    feat = _f32(feat)
    nz, ny, nx, c = feat.shape
    if c != 4:
        raise DropletError("features need 4 channels")
    out = np.empty((nz, ny, nx), dtype=np.float32)
    _check(lib().dg_guess_forward(_i(nx), _i(ny), _i(nz), _pf(feat), _pf(out)), "dg_guess_forward")
    return out


def gradient(feat, target):
    """(loss, flat gradient of 0.5*mean((out-target)^2))."""
    feat = _f32(feat)
    nz, ny, nx, _ = feat.shape
    target = _f32(target, (nz, ny, nx))
    g = np.empty(nparams(), dtype=np.float32)
    loss = ctypes.c_double(0.0)
    _check(lib().dg_guess_gradient(_i(nx), _i(ny), _i(nz), _pf(feat), _pf(target), ctypes.byref(loss), _pf(g)),
           "dg_guess_gradient")
    return loss.value, g


def train_sample(feat, target, lr):
    feat = _f32(feat)
    nz, ny, nx, _ = feat.shape
    target = _f32(target, (nz, ny, nx))
    loss = ctypes.c_double(0.0)
    lr_c = ctypes.c_float(lr)
    _check(lib().dg_guess_train_sample(_i(nx), _i(ny), _i(nz), _pf(feat), _pf(target), ctypes.byref(lr_c),
                                       ctypes.byref(loss)), "dg_guess_train_sample")
    return loss.value


def capture(capacity):
    _check(lib().dg_guess_capture(_i(capacity)), "dg_guess_capture")


def store_count():
    n = ctypes.c_int(0)
    _check(lib().dg_guess_store_count(ctypes.byref(n)), "dg_guess_store_count")
    return n.value


def store_get(i):
    """Sample i of the store: (features [nz, ny, nx, 4], target [nz, ny, nx])."""
    dims = (ctypes.c_int * 3)()
    _check(lib().dg_guess_store_get(_i(i), None, None, dims), "dg_guess_store_get")
    nx, ny, nz = dims
    feat = np.empty((nz, ny, nx, 4), dtype=np.float32)
    tgt = np.empty((nz, ny, nx), dtype=np.float32)
    _check(lib().dg_guess_store_get(_i(i), _pf(feat), _pf(tgt), dims), "dg_guess_store_get")
    return feat, tgt


def train(epochs, lr, seed=0):
    """Adam over the stored samples; returns the mean loss of each epoch."""
    out = np.zeros(epochs, dtype=np.float64)
    lr_c = ctypes.c_float(lr)
    _check(lib().dg_guess_train(_i(epochs), ctypes.byref(lr_c), _i(seed),
                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "dg_guess_train")
    return out


def stats():
    a, r, sk = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    al, mg = ctypes.c_double(0.0), ctypes.c_double(0.0)
    _check(lib().dg_guess_stats(ctypes.byref(a), ctypes.byref(r), ctypes.byref(sk), ctypes.byref(al),
                                ctypes.byref(mg)), "dg_guess_stats")
    return dict(accepted=a.value, rejected=r.value, skipped=sk.value, last_alpha=al.value, mean_gain=mg.value)


def set_controller(probe_every=8, min_gain=4.0):
    """Fallback controller: a learned step dividing |r|^2 by less than
    min_gain sends the next probe_every * 2^(w-1) solves to the zero guess
    alone, w the number of such steps in a row (capped at 16 probe_every)."""
    pe, mg = ctypes.c_int(int(probe_every)), ctypes.c_double(float(min_gain))
    _check(lib().dg_guess_set_controller(ctypes.byref(pe), ctypes.byref(mg)), "dg_guess_set_controller")


def eval_mode(on):
    _check(lib().dg_guess_eval(_i(1 if on else 0)), "dg_guess_eval")


def eval_iters():
    it = (ctypes.c_int * 3)()
    _check(lib().dg_guess_eval_iters(it), "dg_guess_eval_iters")
    return dict(zero=it[0], previous=it[1], learned=it[2])


def release():
    _check(lib().dg_guess_release(), "dg_guess_release")


# ---- single layers ----------------------------------------------------------------

def conv3d_fwd(x, w, bias=None, cout=None, relu=False):
    """x [nz, ny, nx, cinp], w [coutp, 27, cinp] -> y [nz, ny, nx, coutp]
    (the BF16 result; channels >= cout are left zero)."""
    x, w = _f32(x), _f32(w)
    nz, ny, nx, cinp = x.shape
    coutp = w.shape[0]
    cout = coutp if cout is None else cout
    b = None if bias is None else _f32(bias, (coutp,))
    y = np.zeros((nz, ny, nx, coutp), dtype=np.float32)
    _check(lib().dg_conv3d_fwd(_i(nx), _i(ny), _i(nz), _pf(x), _i(cinp), _pf(w), None if b is None else _pf(b),
                               _i(cout), _i(coutp), _i(1 if relu else 0), _pf(y)), "dg_conv3d_fwd")
    return y


def conv3d_dgrad(dy, w):
    """dy [nz, ny, nx, coutp], w [coutp, 27, cinp] -> dx [nz, ny, nx, cinp]."""
    dy, w = _f32(dy), _f32(w)
    nz, ny, nx, coutp = dy.shape
    cinp = w.shape[2]
    dx = np.empty((nz, ny, nx, cinp), dtype=np.float32)
    _check(lib().dg_conv3d_dgrad(_i(nx), _i(ny), _i(nz), _pf(dy), _i(coutp), _pf(w), _i(cinp), _pf(dx)),
           "dg_conv3d_dgrad")
    return dx


def conv3d_wgrad(x, dy):
    """x [.., cinp], dy [.., coutp] -> gw [coutp, 27, cinp], gb [coutp] (FP32)."""
    x, dy = _f32(x), _f32(dy)
    nz, ny, nx, cinp = x.shape
    coutp = dy.shape[3]
    gw = np.empty((coutp, 27, cinp), dtype=np.float32)
    gb = np.empty(coutp, dtype=np.float32)
    _check(lib().dg_conv3d_wgrad(_i(nx), _i(ny), _i(nz), _pf(x), _i(cinp), _pf(dy), _i(coutp), _pf(gw), _pf(gb)),
           "dg_conv3d_wgrad")
    return gw, gb
