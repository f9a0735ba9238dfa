"""The CPU restatement against the golden fixtures written by the
flang-built reference (tests/golden/make_golden.py).  Runs everywhere
(the fixtures travel; the reference does not).  Bar: bit-exact."""
import os

import numpy as np
import pytest

from oracle_ffi import Oracle, OracleKernels as OK

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def test_plic_vectors_bitexact(oracle):
    z = load("plic_vectors.npz")
    m, cc = z["al_m"], z["al_cc"]
    got = np.array([OK.al3d(m[q, 0], m[q, 1], m[q, 2], cc[q]) for q in range(len(cc))])
    assert np.array_equal(got, z["al_out"])
    mm = z["fl_m"]
    got = np.array([OK.fl3d(mm[q, 0], mm[q, 1], mm[q, 2], z["fl_alpha"][q], z["fl_r0"][q], z["fl_dr0"][q])
                    for q in range(len(mm))])
    assert np.array_equal(got, z["fl_out"])


def test_kernels_bitexact(oracle):
    z = load("kernels_18.npz")
    u, rhs, a, b, c, ei = (z[k].copy() for k in ("in_u", "in_rhs", "in_a", "in_b", "in_c", "in_ei"))
    u1 = u.copy()
    OK.relax(u1, 1.0, rhs, a, b, c, ei, 1.5)
    assert np.array_equal(u1, z["relax"])
    res = np.zeros_like(u)
    OK.resid(1.0, res, u1, rhs, a, b, c)
    assert np.array_equal(res, z["resid"])
    uc = np.zeros((10, 10, 10))
    OK.rstrct(uc, u)
    assert np.array_equal(uc, z["rstrct"])
    uf = rhs.copy()
    OK.interp(uf, uc)
    assert np.array_equal(uf, z["interp"])
    g, r2 = u.copy(), np.zeros_like(u)
    OK.addint(g, uc, r2)
    assert np.array_equal(g, z["addint_uf"]) and np.array_equal(r2, z["addint_res"])
    ca = [np.zeros((10, 10, 10)) for _ in range(4)]
    OK.acfromac(*ca, a, b, c)
    for x, k in zip(ca, ("acfromac_a", "acfromac_b", "acfromac_c", "acfromac_ei")):
        assert np.array_equal(x, z[k]), k
    cc = z["in_cc"].copy()
    ab = [np.zeros_like(cc) for _ in range(4)]
    OK.abcei(*ab, cc, 0.4994, 0.0012)
    for x, k in zip(ab, ("abcei_a", "abcei_b", "abcei_c", "abcei_ei")):
        assert np.array_equal(x, z[k]), k
    cf = np.zeros_like(cc)
    OK.filterc27(cc, cf)
    assert np.array_equal(cf, z["filterc27"])
    sc = u.copy()
    OK.bc_scalar(sc)
    assert np.array_equal(sc, z["bc_scalar"])
    u6 = z["in6_u"].copy()
    OK.slvsml(u6, z["in6_rhs"].copy(), z["in6_a"].copy(), z["in6_b"].copy(), z["in6_c"].copy(),
              z["in6_ei"].copy(), 7, 1.5, 0.004)
    assert np.array_equal(u6, z["slvsml"])


def _check_run(z, nsteps, fields, init=None):
    txt = str(z["inputdata"])
    o = Oracle(txt)
    for k, v in (init or {}).items():
        o.field(k)[...] = v
    hist = o.run(nsteps)
    for key in ("tau", "time", "tauh", "tauh2", "npre", "ntop", "npost", "tswap", "vol"):
        got = np.array([h[key] for h in hist])
        assert np.array_equal(got, z[f"h_{key}"]), key
    for f in fields:
        assert np.array_equal(o.field(f), z[f"f_{f}"]), f
    return o


def test_run_34_bitexact(oracle):
    z = load("run_34.npz")
    _check_run(z, 10, ("cc", "u", "v", "w", "p", "s12", "s13", "s23"))


def test_run_66x34x34_bitexact(oracle):
    z = load("run_66x34x34.npz")
    _check_run(z, 6, ("cc", "u", "v", "w", "p"))


def test_run_upwind_keps_bitexact(oracle):
    z = load("run_18_upwind_keps.npz")
    _check_run(z, 5, ("cc", "u", "v", "w", "p", "ktur", "eps", "mut"),
               init=dict(ktur=float(z["init_ktur"]), eps=float(z["init_eps"])))


@pytest.mark.slow
def test_run_66_history_bitexact(oracle):
    z = load("run_66_hist.npz")
    txt = str(z["inputdata"])
    o = Oracle(txt)
    hist = o.run(17)
    for key in ("tau", "ntop", "tswap", "vol"):
        assert np.array_equal(np.array([h[key] for h in hist]), z[f"h_{key}"]), key
    for f in ("cc", "u", "v", "w", "p"):
        inner = o.field(f)[1:-1, 1:-1, 1:-1]
        assert inner.sum() == z[f"sum_{f}"]
        assert (inner * inner).sum() == z[f"sumsq_{f}"]
        assert np.abs(inner).max() == z[f"maxabs_{f}"]


def test_shipped_sgip_calch3d_bitexact(oracle):
    """The reference's shipped parameters: CalcH3d interface heights
    (hswitch=1), SGIP surface tension with its contact-angle wall row, 4
    steps; then SGIP2 and CSF with CalcH3d, 3 steps (checksums).  The
    initial volume fraction is the clipped one the fixture holds (unclipped,
    the reference STOPs at step 0, make_golden.py)."""
    z = load("run_shipped_sgip.npz")
    o = _check_run(z, 4, ("cc", "u", "v", "w", "p", "s11", "s22", "s33", "hxmin", "hxmax", "hymin", "hymax",
                          "hzmin", "hzmax"), init=dict(cc=z["init_cc"]))
    assert o.st.stop_code == 0
    for tag in ("sgip2", "csfh"):
        o = Oracle(str(z[f"{tag}_inputdata"]))
        o.field("cc")[...] = z["init_cc"]
        hist = o.run(3)
        for key in ("tau", "time", "ntop", "tswap", "vol"):
            assert np.array_equal(np.array([h[key] for h in hist]), z[f"{tag}_h_{key}"]), (tag, key)
        for f in ("cc", "u", "v", "w", "p", "hxmin", "hymax", "hzmin"):
            inner = o.field(f)[1:-1, 1:-1, 1:-1]
            assert inner.sum() == z[f"{tag}_sum_{f}"], (tag, f)
            assert (inner ** 2).sum() == z[f"{tag}_sumsq_{f}"], (tag, f)
