"""The remaining turbulence models and the multi-droplet initial condition
on the GPU against the CPU restatement (itself bit-exact against the
flang-built reference on the same cases, test_oracle_reference.py).

- eddyvisc (main.f:1828-1885), the mixing-length model of iturb = 1 and 2:
  whole steps to the whole-step parity bar, 1e-10 relative L2 (measured
  1e-15 to 2e-12 after five steps: the model's own arithmetic is sqrt
  only, the step's other operators differ in the last bits), tau to 1e-12,
  NTOP and tswap equal.
- uv2p (main.f:2032-2155), the anisotropic k-epsilon of iturb = 4: the
  operator from the same state to 1e-13 of each field's maximum (the
  device's acos and cos may differ from the host's in the last bit); whole
  steps (kepsmodel's Reynolds-stress production, convective's Reynolds
  terms in the upwind and central forms) to 1e-8 relative L2 over three
  steps, where the k-epsilon model amplifies last-bit differences (a 1e-12
  relative change of the state grows to 8e-8 in four steps on the CPU
  statement alone).
- the multi-droplet initial condition (dg_init_droplets): one droplet at
  smooth2's centre reproduces dg_init's state bit for bit; three droplets
  give the volume fractions of the sub-point count restated here in numpy,
  bit for bit, and the droplet and gas velocities of the mixture rule.
"""
import numpy as np
import pytest

from oracle_ffi import Oracle, inputdata_text, sampled_cc

pytestmark = pytest.mark.gpu

FIELDS = ("cc", "u", "v", "w", "p", "mut")


def rel_l2(a, b):
    a, b = a[1:-1, 1:-1, 1:-1], b[1:-1, 1:-1, 1:-1]
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


EDDY = [
    ("18^3 iturb 1", dict(nx=18, ny=18, nz=18, iturb=1, xvel1=1.0, yvel1=0.0, y0=0.002)),
    ("34x18x18 iturb 1, inflow 0.2", dict(nx=34, ny=18, nz=18, iturb=1, xvel1=1.0, yvel1=0.0, xvel2=0.2, y0=0.002,
                                          radius=0.0006)),
    ("34x18x18 iturb 2", dict(nx=34, ny=18, nz=18, iturb=2, xvel1=0.8, yvel1=0.0, y0=0.002, radius=0.0005)),
]


@pytest.mark.parametrize("name,kw", EDDY, ids=[e[0] for e in EDDY])
def test_mixing_length_steps_match_restatement(dg, name, kw):
    nsteps = 5
    print(f"case {name}, {nsteps} steps", flush=True)
    txt = inputdata_text(**kw)
    o = Oracle(txt)
    g = dg.DropletGPU(inputdata_text=txt)
    try:
        for t in range(nsteps):
            o.step(t)
            g.step(t)
            st = g.state()
            assert abs(st["tau"] - o.st.tau) <= 1e-12 * o.st.tau
            assert st["tswap"] == o.st.tswap and st["ntop"] == o.st.ntop
        errs = {f: rel_l2(g.get(f), o.field(f)) for f in FIELDS}
        print({k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
        assert np.abs(o.field("mut")).max() > 0.0
        for f, e in errs.items():
            assert e <= 1e-10, (f, e)
    finally:
        g.close()
        o.close()


TURB_INIT = dict(ktur=2e-4, eps=3e-3 * 0.004)
REY = ("cmu", "mut", "u2p", "v2p", "w2p", "uvp", "uwp", "vwp")
STATE = ("cc", "u", "v", "w", "p", "rhon", "mun", "mut", "ktur", "eps", "cmu", "u2p", "v2p", "w2p", "uvp", "uwp",
         "vwp")


def test_uv2p_from_the_same_state(dg):
    print("case uv2p on the restatement's state after 2 steps, 18^3 iturb 4", flush=True)
    txt = inputdata_text(nx=18, ny=18, nz=18, iturb=4)
    o = Oracle(txt)
    for k, v in TURB_INIT.items():
        o.field(k)[...] = v
    o.run(2)
    g = dg.DropletGPU(inputdata_text=txt)
    try:
        for f in STATE:
            g.set(f, o.field(f))
        import oracle_ffi
        oracle_ffi.lib().or_uv2p(o.h)
        g.uv2p()
        worst = {}
        for f in REY:
            b = o.field(f)
            worst[f] = float(np.abs(g.get(f) - b).max() / max(np.abs(b).max(), 1e-300))
        print({k: f"{v:.1e}" for k, v in worst.items()}, flush=True)
        assert np.abs(o.field("uvp")).max() > 0.0
        for f, e in worst.items():
            assert e <= 1e-13, (f, e)
    finally:
        g.close()
        o.close()


ANISO = [
    ("18^3 central", dict(nx=18, ny=18, nz=18, iturb=4), TURB_INIT),
    ("34x18x18 upwind, inflow", dict(nx=34, ny=18, nz=18, iturb=4, convectype=1, xvel2=0.5, y0=0.002, radius=0.0005),
     dict(ktur=1e-4, eps=1e-3 * 0.004)),
]


@pytest.mark.parametrize("name,kw,init", ANISO, ids=[a[0] for a in ANISO])
def test_anisotropic_steps_match_restatement(dg, name, kw, init):
    nsteps = 3
    print(f"case iturb 4 {name}, {nsteps} steps", flush=True)
    txt = inputdata_text(**kw)
    o = Oracle(txt)
    g = dg.DropletGPU(inputdata_text=txt)
    try:
        for k, v in init.items():
            o.field(k)[...] = v
            g.set(k, o.field(k))
        for t in range(nsteps):
            o.step(t)
            g.step(t)
            assert g.state()["tswap"] == o.st.tswap
        errs = {f: rel_l2(g.get(f), o.field(f)) for f in FIELDS + ("ktur", "eps", "u2p", "uvp")}
        print({k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
        assert np.abs(o.field("uvp")).max() > 0.0
        for f, e in errs.items():
            assert e <= 1e-8, (f, e)
    finally:
        g.close()
        o.close()


def test_single_droplet_reproduces_smooth2(dg):
    print("case dg_init_droplets with smooth2's droplet, 34^3", flush=True)
    txt = inputdata_text(nx=34, ny=34, nz=34)
    g = dg.DropletGPU(inputdata_text=txt)
    try:
        ref = {f: g.get(f) for f in ("cc", "u", "v", "w", "rhon", "mun")}
        p = g.smooth2_droplet()
        g.init_droplets(p[None, :])
        for f, a in ref.items():
            assert np.array_equal(g.get(f), a), f
    finally:
        g.close()


def test_three_droplets_match_sampling(dg):
    print("case three droplets on 34^3", flush=True)
    txt = inputdata_text(nx=34, ny=34, nz=34)
    g = dg.DropletGPU(inputdata_text=txt)
    try:
        p0 = g.smooth2_droplet()
        h = 1.0 / 32.0
        drops = np.array([[0.30, 0.30, 0.50, 0.12, 0.0, -1.0, 0.0],
                          [0.65, 0.55, 0.40, 0.15, 0.2, -0.5, 0.1],
                          [0.50, 0.80, 0.70, 0.10, 0.0, -2.0, 0.0]])
        g.init_droplets(drops)
        cc = g.get("cc")
        want = sampled_cc(cc.shape, h, drops)
        assert np.array_equal(cc[1:-1, 1:-1, 1:-1], want[1:-1, 1:-1, 1:-1])
        vol = cc[1:-1, 1:-1, 1:-1].sum() * h ** 3
        exact = (4.0 / 3.0) * np.pi * (drops[:, 3] ** 3).sum()
        assert abs(vol - exact) / exact < 2e-3, (vol, exact)
        v = g.get("v")
        # a full cell of each droplet moves with it; an empty cell with the gas
        for x0, y0, z0, r, uu, vv, ww in drops:
            # the cell around the centre: cell i spans [(i - 2.5) h, (i - 1.5) h]
            i, j, k = (int(round(c / h + 2.0)) for c in (x0, y0, z0))
            assert cc[k - 1, j - 1, i - 1] == 1.0 and v[k - 1, j - 1, i - 1] == vv
        st_yvel2 = 0.0
        assert v[2, 2, 2] == st_yvel2 and cc[2, 2, 2] == 0.0
        assert p0.shape == (7,)
    finally:
        g.close()


def test_droplet_parameters_rejected(dg):
    print("case invalid droplet parameters", flush=True)
    g = dg.DropletGPU(inputdata_text=inputdata_text(nx=18, ny=18, nz=18))
    try:
        with pytest.raises(dg.DropletError, match="radius"):
            g.init_droplets(np.array([[0.5, 0.5, 0.5, -0.1, 0, 0, 0]]))
        many = np.array([[0.5, 0.5, 0.5, 0.2 + 0.001 * q, 0, 0, 0] for q in range(9)])
        with pytest.raises(dg.DropletError, match="more than 8 droplets"):
            g.init_droplets(many)
    finally:
        g.close()
