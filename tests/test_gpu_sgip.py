"""The reference's shipped configuration on the GPU: CalcH3d interface
heights (hswitch = 1, main.f:1312-1810) and SGIP surface tension with its
contact-angle wall row (surfacet = 5, main.f:2835-3027), plus SGIP2
(surfacet = 6, main.f:3029-3166) and CSF with CalcH3d, against the CPU
restatement (itself bit-exact with the flang-built reference on these
inputs, test_oracle_reference.py, and with the committed fixture,
test_oracle_golden.py).

Parameters are those of the reference's inputdata.in: 66x34x34, densities
1000/500, inflow 0.5 m/s, no gravity, a 1 mm droplet centred 0.17 mm below
the wall, the reference's multigrid schedule (NPRE 5, NTOP 4, NPOST 5,
omega 1.5).  Both sides start from the initial volume fraction clipped to
[0, 1]: unclipped, the reference itself STOPs at step 0 (smooth2's 1 + 4e-13
cells turn to NaN in AL3D and CalcH3d finds no case for a NaN normal).

Two bars:
- The operators from the same state (the restatement's whole state copied
  into the GPU): CalcH3d's heights, SGIP's and SGIP2's S11..S33 and w11..w33
  on every cell, ghosts included, to 1e-12 of the field's maximum (the
  differences are the last bits of the device's sin/cos in SGIP's
  contact-angle row and pow in CalcH3d's triangle case).
- Whole steps.  CalcH3d picks its case with exact ties (absxn == n1,
  |absyn - n2| < 1e-6; main.f:1602-1728), and the droplet's symmetry puts
  many cells on a tie, so a last-bit difference of the volume fraction (the
  device's pow/acos/cos in AL3D) can move a cell's heights by O(1); with
  SGIP the heights feed the flow.  The restatement shows the same
  sensitivity against itself: one ulp added to every interface cell of the
  initial volume fraction moves its fields after 4 steps by up to 1.5e-4
  (SGIP) and 3.9e-4 (SGIP2) relative L2, and by 1e-14 with CSF + CalcH3d,
  where the heights do not feed back.  The bar: each field within 10 times
  that self-sensitivity (measured in the test) plus 1e-10; tau to 1e-12,
  NTOP and tswap equal every step.
"""
import numpy as np
import pytest

from oracle_ffi import Oracle, clipped_cc, shipped_text

pytestmark = pytest.mark.gpu

HEIGHTS = ("hxmin", "hxmax", "hymin", "hymax", "hzmin", "hzmax")
FIELDS = ("cc", "u", "v", "w", "p", "s11", "s22", "s33") + HEIGHTS


def rel_l2(a, b):
    a, b = a[1:-1, 1:-1, 1:-1], b[1:-1, 1:-1, 1:-1]
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def pair(dg, txt):
    cc0 = clipped_cc(txt)
    o = Oracle(txt)
    o.field("cc")[...] = cc0
    g = dg.DropletGPU(inputdata_text=txt)
    g.set("cc", cc0)
    return o, g


def oracle_self_sensitivity(txt, nsteps):
    """Relative L2 between two restatement runs whose initial volume
    fractions differ by one ulp in every interface cell."""
    cc0 = clipped_cc(txt)
    out = []
    for bump in (False, True):
        o = Oracle(txt)
        c = cc0.copy()
        if bump:
            m = (c > 0) & (c < 1)
            c[m] = np.nextafter(c[m], 2.0)
        o.field("cc")[...] = c
        o.run(nsteps)
        out.append({f: o.field(f).copy() for f in FIELDS})
        o.close()
    return {f: rel_l2(out[1][f], out[0][f]) for f in FIELDS}


@pytest.mark.parametrize("surfacet,nsteps", [(5, 4), (6, 3), (2, 3)], ids=["SGIP", "SGIP2", "CSF+CalcH3d"])
def test_shipped_configuration_matches_restatement(dg, surfacet, nsteps):
    print(f"case shipped 66x34x34 hswitch=1 surfacet={surfacet}, {nsteps} steps", flush=True)
    txt = shipped_text(surfacet=surfacet)
    sens = oracle_self_sensitivity(txt, nsteps)
    o, g = pair(dg, txt)
    try:
        for t in range(nsteps):
            o.step(t)
            g.step(t)
            assert o.st.stop_code == 0
            st = g.state()
            assert abs(st["tau"] - o.st.tau) <= 1e-12 * o.st.tau, (t, st["tau"], o.st.tau)
            assert st["ntop"] == o.st.ntop and st["tswap"] == o.st.tswap
        errs = {f: rel_l2(g.get(f), o.field(f)) for f in FIELDS}
        print("GPU vs restatement", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
        print("restatement vs itself (+1 ulp)", {k: f"{v:.1e}" for k, v in sens.items()}, flush=True)
        for f, e in errs.items():
            assert e <= 10.0 * sens[f] + 1e-10, (f, e, sens[f])
    finally:
        g.close()
        o.close()


STATE = ("cc", "u", "v", "w", "p", "s11", "s22", "s33", "s12", "s13", "s23", "rhon", "mun", "res", "rhs", "a",
         "b", "c", "ei") + HEIGHTS


@pytest.mark.parametrize("surfacet", [5, 6], ids=["SGIP", "SGIP2"])
def test_operators_from_the_same_state(dg, surfacet):
    """After a step and the next VOFmethod on the restatement, its whole
    state is copied into the GPU; CalcH3d, then SGIP or SGIP2, on both."""
    import oracle_ffi
    print(f"case shipped 66x34x34 CalcH3d + surfacet={surfacet} from the same state", flush=True)
    txt = shipped_text(surfacet=surfacet)
    o, g = pair(dg, txt)
    L = oracle_ffi.lib()
    try:
        o.step(0)
        L.or_vofmethod(o.h, 1)
        L.or_checktime(o.h, 1)
        for f in STATE:
            g.set(f, o.field(f))
        g.set_state(tau=o.st.tau, time=o.st.time, tauh=o.st.tauh, tauh2=o.st.tauh2, tswap=o.st.tswap,
                    npre=o.st.npre, ntop=o.st.ntop, npost=o.st.npost)
        L.or_calch3d(o.h)
        g.calch3d()
        assert o.st.stop_code == 0
        worst = {f: float(np.abs(g.get(f) - o.field(f)).max()) for f in HEIGHTS}
        print("heights", {k: f"{v:.1e}" for k, v in worst.items()}, flush=True)
        for f, e in worst.items():
            assert e <= 1e-12, (f, e)
        if surfacet == 5:
            L.or_sgip(o.h)
            g.sgip()
        else:
            L.or_sgip2(o.h)
            g.sgip2()
        worst = {}
        for f in ("s11", "s22", "s33", "p", "res", "rhs"):
            a, b = g.get(f), o.field(f)
            worst[f] = float(np.abs(a - b).max()) / max(float(np.abs(b).max()), 1e-300)
        print("stresses", {k: f"{v:.1e}" for k, v in worst.items()}, flush=True)
        for f, e in worst.items():
            assert e <= 1e-12, (f, e)
    finally:
        g.close()
        o.close()
