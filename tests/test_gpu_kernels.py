"""GPU parity of the reference's explicit-argument kernels against the CPU
restatement (oracle), through the C-ABI host-pointer entry points.

Bar: bit-exact for the multigrid/BC/filter kernels (the same FP64 operations
in the same order, no FMA on either side); AL3D/FL3D within 4 ulp-scale
relative error because they call pow/acos/cos/sqrt (device OCML versus glibc).
"""
import numpy as np
import pytest

from oracle_ffi import OracleKernels as OK

pytestmark = pytest.mark.gpu

SHAPES = [(18, 18, 18), (34, 18, 18), (10, 6, 6)]


def _rand(shape, rng, lo=-1.0, hi=1.0):
    nx, ny, nz = shape
    return rng.uniform(lo, hi, size=(nz, ny, nx))


def _coef(shape, rng):
    # face inverse densities between 1 and 833 like a water/air split
    return np.where(rng.uniform(size=(shape[2], shape[1], shape[0])) < 0.5, 1.0, 833.3333) * rng.uniform(
        0.9, 1.1, size=(shape[2], shape[1], shape[0]))


@pytest.mark.parametrize("shape", SHAPES)
def test_relax_resid_bitexact(dg, shape):
    print("case relax/resid", shape, flush=True)
    rng = np.random.default_rng(1)
    u, rhs = _rand(shape, rng), _rand(shape, rng)
    a, b, c = _coef(shape, rng), _coef(shape, rng), _coef(shape, rng)
    ei = 1.0 / (a + b + c)
    u1, u2 = u.copy(), u.copy()
    OK.relax(u1, 1.0, rhs, a, b, c, ei, 1.5)
    dg.relax(u2, 1.0, rhs, a, b, c, ei, 1.5)
    assert np.array_equal(u1, u2)
    r1, r2 = np.zeros_like(u), np.zeros_like(u)
    OK.resid(1.0, r1, u1, rhs, a, b, c)
    dg.resid(1.0, r2, u2, rhs, a, b, c)
    assert np.array_equal(r1, r2)


@pytest.mark.parametrize("shape", SHAPES)
def test_transfers_bitexact(dg, shape):
    print("case transfers", shape, flush=True)
    rng = np.random.default_rng(2)
    nx, ny, nz = shape
    cshape = (nx // 2 + 1, ny // 2 + 1, nz // 2 + 1)
    uf = _rand(shape, rng)
    uc1, uc2 = np.zeros((cshape[2], cshape[1], cshape[0])), np.zeros((cshape[2], cshape[1], cshape[0]))
    OK.rstrct(uc1, uf)
    dg.rstrct(uc2, uf)
    assert np.array_equal(uc1, uc2)
    f1, f2 = _rand(shape, rng), None
    f2 = f1.copy()
    OK.interp(f1, uc1)
    dg.interp(f2, uc2)
    assert np.array_equal(f1, f2)
    g1, res1 = _rand(shape, rng), np.zeros(uf.shape)
    g2, res2 = g1.copy(), res1.copy()
    OK.addint(g1, uc1, res1)
    dg.addint(g2, uc2, res2)
    assert np.array_equal(g1, g2) and np.array_equal(res1, res2)
    a2, b2, c2 = _coef(shape, rng), _coef(shape, rng), _coef(shape, rng)
    outs1 = [np.zeros_like(uc1) for _ in range(4)]
    outs2 = [np.zeros_like(uc1) for _ in range(4)]
    OK.acfromac(*outs1, a2, b2, c2)
    dg.acfromac(*outs2, a2, b2, c2)
    for x, y in zip(outs1, outs2):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("shape", SHAPES)
def test_abcei_filter_bcscalar_bitexact(dg, shape):
    print("case abcei/filter/bc_scalar", shape, flush=True)
    rng = np.random.default_rng(3)
    cc = rng.uniform(0.0, 1.0, size=(shape[2], shape[1], shape[0]))
    cc[rng.uniform(size=cc.shape) < 0.3] = 1.0
    o1 = [np.zeros_like(cc) for _ in range(4)]
    o2 = [np.zeros_like(cc) for _ in range(4)]
    OK.abcei(*o1, cc, 0.4994, 0.0012)
    dg.abcei(*o2, cc, 0.4994, 0.0012)
    for x, y in zip(o1, o2):
        assert np.array_equal(x, y)
    f1, f2 = np.zeros_like(cc), np.zeros_like(cc)
    OK.filterc27(cc, f1)
    dg.filterc27(cc, f2)
    assert np.array_equal(f1, f2)
    s1 = _rand(shape, rng)
    s2 = s1.copy()
    OK.bc_scalar(s1)
    dg.bc_scalar(s2)
    assert np.array_equal(s1, s2)


def test_slvsml_bitexact(dg):
    print("case slvsml 6x4x4", flush=True)
    rng = np.random.default_rng(4)
    shape = (6, 4, 4)
    u, rhs = _rand(shape, rng), _rand(shape, rng)
    a, b, c = _coef(shape, rng), _coef(shape, rng), _coef(shape, rng)
    ei = 1.0 / (a + b + c)
    u1, u2 = u.copy(), u.copy()
    OK.slvsml(u1, rhs, a, b, c, ei, 7, 1.5, 0.004)
    dg.slvsml(u2, rhs, a, b, c, ei, 7, 1.5, 0.004)
    assert np.array_equal(u1, u2)


def test_plic_functions(dg):
    print("case AL3D/FL3D batches", flush=True)
    rng = np.random.default_rng(5)
    n = 20000
    m = rng.uniform(0, 1, size=(n, 3))
    m /= m.sum(axis=1, keepdims=True)
    cc = rng.uniform(0, 1, size=n)
    cc[:200] = rng.uniform(0, 1e-6, size=200)
    cc[200:400] = 1 - rng.uniform(0, 1e-6, size=200)
    got = dg.al3d(m[:, 0], m[:, 1], m[:, 2], cc)
    ref = np.array([OK.al3d(*m[q], cc[q]) for q in range(n)])
    np.testing.assert_allclose(got, ref, rtol=2e-14, atol=1e-15)
    mm = rng.uniform(-1, 1, size=(n, 3))
    al = rng.uniform(-0.5, 1.5, size=n)
    r0 = rng.uniform(-0.3, 1.0, size=n)
    dr0 = rng.uniform(0.01, 1.0, size=n)
    got = dg.fl3d(mm[:, 0], mm[:, 1], mm[:, 2], al, r0, dr0)
    ref = np.array([OK.fl3d(*mm[q], al[q], r0[q], dr0[q]) for q in range(n)])
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-15)


def test_fortran_names_match_oracle(dg):
    """The fine-grained routines under the reference's own symbols (relax_,
    resid_, rstrct_, abcei_, filterc27_, al3d_, fl3d_), called by reference
    as a Fortran program calls them, against the restatement: bit-exact, and
    AL3D/FL3D to the same 2e-14 as their batch entry points."""
    import ctypes
    print("case Fortran-named kernels 18x18x18", flush=True)
    L = dg.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    P = lambda a: a.ctypes.data_as(dp)  # noqa: E731
    I = lambda v: ctypes.byref(ctypes.c_int(v))  # noqa: E731
    D = lambda v: ctypes.byref(ctypes.c_double(v))  # noqa: E731
    L.dg_set_stop_on_error_(I(0))
    shape = (18, 18, 18)
    rng = np.random.default_rng(11)
    u, rhs = _rand(shape, rng), _rand(shape, rng)
    a, b, c = _coef(shape, rng), _coef(shape, rng), _coef(shape, rng)
    ei = 1.0 / (a + b + c)
    u1, u2 = u.copy(), u.copy()
    OK.relax(u1, 1.0, rhs, a, b, c, ei, 1.5)
    L.relax_(P(u2), D(1.0), P(rhs), P(a), P(b), P(c), P(ei), I(2), I(18), I(18), I(18), D(1.5))
    assert np.array_equal(u1, u2)
    r1, r2 = np.zeros_like(u), np.zeros_like(u)
    OK.resid(1.0, r1, u1, rhs, a, b, c)
    L.resid_(D(1.0), P(r2), P(u2), P(rhs), P(a), P(b), P(c), I(2), I(18), I(18), I(18))
    assert np.array_equal(r1, r2)
    uc1, uc2 = np.zeros((10, 10, 10)), np.zeros((10, 10, 10))
    OK.rstrct(uc1, r1)
    L.rstrct_(P(uc2), P(r2), I(2), I(10), I(10), I(10))
    assert np.array_equal(uc1, uc2)
    cc = rng.uniform(0.0, 1.0, size=(18, 18, 18))
    f1, f2 = np.zeros_like(cc), np.zeros_like(cc)
    OK.filterc27(cc, f1)
    L.filterc27_(P(cc), P(f2), I(18), I(18), I(18))
    assert np.array_equal(f1, f2)
    outs = [np.zeros_like(cc) for _ in range(8)]
    OK.abcei(*outs[:4], cc, 0.4994, 0.0012)
    L.abcei_(*(P(x) for x in outs[4:]), P(cc), D(0.4994), D(0.0012), I(2), I(18), I(18), I(18))
    for q in range(4):
        assert np.array_equal(outs[q], outs[4 + q])
    import oracle_ffi
    O = oracle_ffi.lib()
    L.al3d_.restype = ctypes.c_double
    L.fl3d_.restype = ctypes.c_double
    for m1, m2, m3, cv in ((0.1, 0.3, 0.6, 0.2), (0.2, 0.2, 0.6, 0.7), (0.0, 0.4, 0.6, 0.05)):
        al_g = L.al3d_(D(m1), D(m2), D(m3), D(cv))
        al_o = O.or_al3d(m1, m2, m3, cv)
        assert abs(al_g - al_o) <= 2e-14 * abs(al_o), (al_g, al_o)
        v_g = L.fl3d_(D(m1), D(m2), D(m3), D(al_o), D(0.0), D(1.0))
        v_o = O.or_fl3d(m1, m2, m3, al_o, 0.0, 1.0)
        assert abs(v_g - v_o) <= 2e-14 * max(abs(v_o), 1e-300), (v_g, v_o)
    L.dg_set_stop_on_error_(I(1))
