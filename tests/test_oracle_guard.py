"""The VOF guard (a deviation from the reference, off by default): a partly
filled cell whose Youngs normal vanishes is advected with a plane normal to
the sweep.  The reference evaluates 0/0 in AL3D/FL3D there (swp,
main.f:848-1000) and the run turns to NaN.

Pinned by: (1) the reference itself going NaN on the case the guard exists
for, with the unguarded restatement bit-exact to it; (2) the guard changing
nothing while no such cell exists (bit-identical steps before the trigger);
(3) the closed form of a plane interface moved by a uniform velocity: an
isolated cell of fraction c, fluid on the upwind side of a plane normal to x,
shifted by a (in cell widths) sends max(0, c + a - 1) downstream.
"""
import numpy as np
import pytest

import refrun
from oracle_ffi import Oracle, inputdata_text

NAN_CASE = dict(nx=34, ny=18, nz=18, convectype=1, y0=0.002)


def test_reference_goes_nan_without_guard():
    if not refrun.have_ref(34, 18, 18):
        pytest.skip("reference library for 34x18x18 not built")
    txt = inputdata_text(**NAN_CASE)
    ref = refrun.run_reference(txt, 5, want_fields=("cc", "u"))
    o = Oracle(txt)
    o.run(5)
    assert np.isnan(ref["fields"]["cc"]).any()
    assert np.array_equal(ref["fields"]["cc"], o.field("cc"), equal_nan=True)
    o.close()


def test_guard_identical_before_trigger_and_finite_after():
    txt = inputdata_text(**NAN_CASE)
    a, b = Oracle(txt), Oracle(txt)
    b.st.vof_guard = 1
    for t in range(4):  # the unguarded run first turns NaN at step 4
        a.step(t)
        b.step(t)
        for f in ("cc", "u", "v", "w", "p"):
            assert np.array_equal(a.field(f), b.field(f)), (t, f)
    a.step(4)
    b.step(4)
    assert np.isnan(a.field("cc")).any()
    for t in range(5, 8):
        b.step(t)
    for f in ("cc", "u", "v", "w", "p"):
        assert np.isfinite(b.field(f)).all(), f
    a.close()
    b.close()


@pytest.mark.parametrize("c0,shift", [(0.3, 0.8), (0.3, 0.5), (0.75, 0.4)])
def test_isolated_cell_plane_advection(c0, shift):
    """One partly filled cell in an empty 18^3 box moved along x by a
    uniform velocity: the downstream neighbour receives max(0, c0+a-1), the
    cell keeps the rest, and the other sweeps (zero velocity) move nothing."""
    o = Oracle(inputdata_text(nx=18, ny=18, nz=18))
    o.st.vof_guard = 1
    cc = o.field("cc")
    cc[...] = 0.0
    i, j, k = 8, 9, 9  # 0-based (x, y, z) of the cell, well inside
    cc[k, j, i] = c0
    for f in ("u", "v", "w"):
        o.field(f)[...] = 0.0
    o.field("u")[...] = shift / o.st.tauh
    from oracle_ffi import lib
    import ctypes
    lib().or_vofmethod.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib().or_vofmethod(o.h, 0)
    out = o.field("cc")
    moved = max(0.0, c0 + shift - 1.0)
    assert out[k, j, i + 1] == pytest.approx(moved, abs=1e-14)
    assert out[k, j, i] == pytest.approx(c0 - moved, abs=1e-14)
    mask = np.ones_like(out, dtype=bool)
    mask[k, j, i] = mask[k, j, i + 1] = False
    assert np.abs(out[1:-1, 1:-1, 1:-1][mask[1:-1, 1:-1, 1:-1]]).max() == 0.0
    o.close()


def test_isolated_cell_without_guard_loses_its_fluid():
    o = Oracle(inputdata_text(nx=18, ny=18, nz=18))
    cc = o.field("cc")
    cc[...] = 0.0
    cc[9, 9, 8] = 0.3
    for f in ("u", "v", "w"):
        o.field(f)[...] = 0.0
    o.field("u")[...] = 0.5 / o.st.tauh
    from oracle_ffi import lib
    import ctypes
    lib().or_vofmethod.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib().or_vofmethod(o.h, 0)
    # the 0/0 lands in the cell's own fluxes: depending on which sweep meets
    # it the fraction turns NaN or is clipped away; either way the cell's
    # fluid is not conserved
    vol = float(o.field("cc")[1:-1, 1:-1, 1:-1].sum())
    assert not (abs(vol - 0.3) < 1e-12)
    o.close()
