"""The CPU restatement against the live flang-built reference
(oracle/_ref/libref_*.so from oracle/build_ref.sh).  Skipped where the
reference was not built (machines without /root/reference).  Bar: bit-exact."""
import numpy as np
import pytest

import refrun
from oracle_ffi import Oracle, OracleKernels as OK, inputdata_text


def need(nx, ny, nz):
    if not refrun.have_ref(nx, ny, nz):
        pytest.skip(f"reference library for {nx}x{ny}x{nz} not built")


RUNS = [
    ("18^3 central 8 steps", dict(nx=18, ny=18, nz=18), 8, None),
    ("34^3 central 20 steps", dict(nx=34, ny=34, nz=34), 20, None),
    ("34x18x18 upwind", dict(nx=34, ny=18, nz=18, convectype=1, y0=0.002, radius=0.0005), 10, None),
    ("18^3 no convection, density ratio 2", dict(nx=18, ny=18, nz=18, convectype=0, rho2=500.0), 10, None),
    ("18^3 inflow xvel2=0.5, gravity on x", dict(nx=18, ny=18, nz=18, xvel2=0.5, grav1=-2.0), 6, None),
    ("18^3 more sweeps, omega 1.2", dict(nx=18, ny=18, nz=18, npre=4, ntop=6, npost=3, omega=1.2, cycles=3), 6, None),
    ("18^3 k-epsilon", dict(nx=18, ny=18, nz=18, iturb=3), 5, dict(ktur=2e-4, eps=3e-3 * 0.004)),
    ("34^3 surface tension off", dict(nx=34, ny=34, nz=34, sigma=0.0), 4, None),
    ("18^3 mixing length (iturb 1), liquid moving in x", dict(nx=18, ny=18, nz=18, iturb=1, xvel1=1.0, yvel1=0.0,
                                                             y0=0.002), 5, None),
    ("34x18x18 mixing length (iturb 1), inflow 0.2", dict(nx=34, ny=18, nz=18, iturb=1, xvel1=1.0, yvel1=0.0,
                                                         xvel2=0.2, y0=0.002, radius=0.0006), 5, None),
    ("34x18x18 mixing length (iturb 2), faster drop", dict(nx=34, ny=18, nz=18, iturb=2, xvel1=0.8, yvel1=0.0,
                                                          y0=0.002, radius=0.0005), 5, None),
    ("18^3 anisotropic k-epsilon (iturb 4)", dict(nx=18, ny=18, nz=18, iturb=4), 5, dict(ktur=2e-4, eps=3e-3 * 0.004)),
    ("34x18x18 anisotropic k-epsilon (iturb 4), upwind, inflow",
     dict(nx=34, ny=18, nz=18, iturb=4, convectype=1, xvel2=0.5, y0=0.002, radius=0.0005), 4,
     dict(ktur=1e-4, eps=1e-3 * 0.004)),
]
TURB_FIELDS = ("cmu", "u2p", "v2p", "w2p", "uvp", "uwp", "vwp")


@pytest.mark.parametrize("name,kw,nsteps,init", RUNS, ids=[r[0] for r in RUNS])
def test_whole_steps_bitexact(name, kw, nsteps, init):
    nx, ny, nz = kw["nx"], kw["ny"], kw["nz"]
    need(nx, ny, nz)
    txt = inputdata_text(**kw)
    fields = refrun.FIELD_NAMES + (TURB_FIELDS if kw.get("iturb", 0) else ())
    ref = refrun.run_reference(txt, nsteps, init_fields=init, want_fields=fields)
    o = Oracle(txt)
    for k, v in (init or {}).items():
        o.field(k)[...] = v
    hist = o.run(nsteps)
    for hr, ho in zip(ref["hist"], hist):
        for key in ("tau", "time", "tauh", "tauh2", "npre", "ntop", "npost", "tswap", "vol"):
            assert hr[key] == ho[key], (hr["t"], key, hr[key], ho[key])
    for f in fields:
        assert np.array_equal(ref["fields"][f], o.field(f), equal_nan=True), f
    if kw.get("iturb", 0):
        # the model did something: a nonzero eddy viscosity in the interior
        assert np.abs(o.field("mut")[1:-1, 1:-1, 1:-1]).max() > 0.0
    if kw.get("iturb") == 4:
        assert np.abs(o.field("uvp")[1:-1, 1:-1, 1:-1]).max() > 0.0


def test_nan_from_smooth2_overshoot_reproduced():
    """At 66^3 with r = 0.8 mm the reference's smooth2 leaves cc = 1+4e-13 in
    16 cells (27000 additions of 1/27000) and AL3D then takes a cube root of
    a negative number: the step goes NaN.  The restatement reproduces it."""
    need(66, 66, 66)
    txt = inputdata_text(nx=66, ny=66, nz=66)
    o = Oracle(txt)
    assert int((o.field("cc") > 1.0).sum()) == 16
    ref = refrun.run_reference(txt, 1, want_fields=("cc", "u"))
    o.run(1)
    assert np.array_equal(ref["fields"]["cc"], o.field("cc"), equal_nan=True)
    assert np.isnan(o.field("cc")).sum() > 0
    assert np.array_equal(np.isnan(ref["fields"]["u"]), np.isnan(o.field("u")))


KSHAPES = [(18, 18, 18), (34, 18, 18)]


@pytest.mark.parametrize("shape", KSHAPES)
def test_kernels_random_bitexact(shape):
    need(*shape)
    K = refrun.RefKernels(*shape)
    rng = np.random.default_rng(11)
    nx, ny, nz = shape
    s = (nz, ny, nx)
    u, rhs = rng.uniform(-1, 1, s), rng.uniform(-1, 1, s)
    a, b, c = (np.where(rng.uniform(size=s) < 0.5, 1.0, 833.3) * rng.uniform(0.9, 1.1, s) for _ in range(3))
    ei = 1.0 / (a + b + c)
    u1, u2 = u.copy(), u.copy()
    K.relax(u1, 1.0, rhs, a, b, c, ei, 1.5)
    OK.relax(u2, 1.0, rhs, a, b, c, ei, 1.5)
    assert np.array_equal(u1, u2)
    r1, r2 = np.zeros(s), np.zeros(s)
    K.resid(1.0, r1, u1, rhs, a, b, c)
    OK.resid(1.0, r2, u2, rhs, a, b, c)
    assert np.array_equal(r1, r2)
    cs = (nz // 2 + 1, ny // 2 + 1, nx // 2 + 1)
    c1, c2 = np.zeros(cs), np.zeros(cs)
    K.rstrct(c1, u)
    OK.rstrct(c2, u)
    assert np.array_equal(c1, c2)
    o1 = [np.zeros(cs) for _ in range(4)]
    o2 = [np.zeros(cs) for _ in range(4)]
    K.acfromac(*o1, a, b, c)
    OK.acfromac(*o2, a, b, c)
    for x, y in zip(o1, o2):
        assert np.array_equal(x, y)
    f1, f2 = u.copy(), u.copy()
    r1b, r2b = np.zeros(s), np.zeros(s)
    K.addint(f1, c1, r1b)
    OK.addint(f2, c2, r2b)
    assert np.array_equal(f1, f2) and np.array_equal(r1b, r2b)


def test_shipped_sgip_sgip2_calch3d_bitexact():
    """CalcH3d (main.f:1312-1810), SGIP (main.f:2835-3027, the contact-angle
    wall row with SIND/COSD included) and SGIP2 (main.f:3029-3166) against
    the live reference on its shipped parameters, initial cc clipped to
    [0, 1] on both sides (unclipped the reference STOPs at step 0).  Every
    field and the six height arrays, bit for bit."""
    need(66, 34, 34)
    from oracle_ffi import clipped_cc, shipped_text
    for surfacet, nsteps in ((5, 3), (6, 2)):
        txt = shipped_text(surfacet=surfacet)
        cc0 = clipped_cc(txt)
        ref = refrun.run_reference(txt, nsteps, init_fields={"cc": cc0},
                                   want_fields=refrun.FIELD_NAMES + refrun.HEIGHT_NAMES)
        o = Oracle(txt)
        o.field("cc")[...] = cc0
        hist = o.run(nsteps)
        assert o.st.stop_code == 0
        for hr, ho in zip(ref["hist"], hist):
            for key in ("tau", "time", "npre", "ntop", "npost", "tswap", "vol"):
                assert hr[key] == ho[key], (surfacet, hr["t"], key)
        for f in refrun.FIELD_NAMES + refrun.HEIGHT_NAMES:
            assert np.array_equal(ref["fields"][f], o.field(f), equal_nan=True), (surfacet, f)
        o.close()
