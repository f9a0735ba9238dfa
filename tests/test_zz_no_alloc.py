"""dg_cfl_noalloc (allocates nothing; partials in a fixed device array)
against dg_cfl on the same velocity fields, bit for bit: two calls in a row,
the second on a smaller grid, so the second also shows that the first
call's leftovers in the shared array do not reach its result; plus the
host-side rejection of a grid above the maximum (no kernel is launched)."""
import ctypes

import numpy as np
import pytest

from oracle_ffi import inputdata_text

pytestmark = pytest.mark.gpu


def _hip(dg):
    """The HIP runtime the library itself loaded (libamdhip64.so.7, the same
    object), for the test's own device buffers: the caller's inputs and
    output, which the entry point takes as device pointers."""
    dg.lib()
    H = ctypes.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipFree.argtypes = [ctypes.c_void_p]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    H.hipDeviceSynchronize.argtypes = []
    return H


def _noalloc(dg, u, v, w, tauh):
    H = _hip(dg)
    L = dg.lib()
    nz, ny, nx = u.shape
    nbytes = u.size * 8
    bufs = []
    try:
        for a in (u, v, w):
            p = ctypes.c_void_p()
            assert H.hipMalloc(ctypes.byref(p), nbytes) == 0
            bufs.append(p)
            a = np.ascontiguousarray(a, dtype=np.float64)
            assert H.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), nbytes, 1) == 0  # host to device
        out = ctypes.c_void_p()
        assert H.hipMalloc(ctypes.byref(out), 8) == 0
        bufs.append(out)
        I = lambda x: ctypes.byref(ctypes.c_int(x))  # noqa: E731
        rc = L.dg_cfl_noalloc(bufs[0], bufs[1], bufs[2], I(nx), I(ny), I(nz), ctypes.byref(ctypes.c_double(tauh)),
                              out, None)
        assert rc == 0, L.dg_last_error()
        assert H.hipDeviceSynchronize() == 0
        res = np.zeros(1)
        assert H.hipMemcpy(res.ctypes.data_as(ctypes.c_void_p), out, 8, 2) == 0  # device to host
        return float(res[0])
    finally:
        for p in bufs:
            H.hipFree(p)


@pytest.mark.parametrize("n1,n2", [(34, 18)])
def test_noalloc_cfl_matches_dg_cfl_twice(dg, n1, n2):
    print("case dg_cfl_noalloc vs dg_cfl at", n1, "then", n2, flush=True)
    for n in (n1, n2):
        g = dg.DropletGPU(inputdata_text=inputdata_text(nx=n, ny=n, nz=n))
        try:
            g.step(0)  # a non-trivial velocity field
            rng = np.random.default_rng(n)
            u = g.get("u") + rng.standard_normal((n, n, n)) * 1e-3
            g.set("u", u)
            ref = g.cfl()
            tauh = g.state()["tauh"]
            got = _noalloc(dg, g.get("u"), g.get("v"), g.get("w"), tauh)
            print(n, ref, got, flush=True)
            assert got == ref, (n, got, ref)
        finally:
            g.close()


def test_noalloc_cfl_rejects_oversized_grid(dg):
    L = dg.lib()
    I = lambda x: ctypes.byref(ctypes.c_int(x))  # noqa: E731
    dp = ctypes.c_void_p
    rc = L.dg_cfl_noalloc(dp(0), dp(0), dp(0), I(2048), I(2048), I(1024), ctypes.byref(ctypes.c_double(1.0)),
                          dp(0), None)
    assert rc == -1
    assert b"exceed the maximum" in L.dg_last_error()
