"""Failure detection on the GPU path, forced from the host.

The reference has no error reporting: a NaN velocity passes checktime's two
tests (main.f:2156-2175) and the run carries on, and its multigrid schedule
runs a fixed number of cycles whatever the residual (main.f:3645-3763).
Here a non-finite CFL, a non-finite or unconverged tolerance-mode solve and
a failed coarsest factorisation make the entry point return -1 with an
error message (DropletError in Python), and the solver's diagnostics stay
readable.  Every condition is set up on the host (a NaN written with
dg_set_field, an iteration cap of 1 against a 1e-14 tolerance), so every
kernel still runs to its normal end.
"""
import numpy as np
import pytest

from oracle_ffi import inputdata_text

pytestmark = pytest.mark.gpu


def _seeded_system(n, seed):
    rng = np.random.default_rng(seed)
    shape = (n, n, n)
    rhs = rng.standard_normal(shape)
    cc = (rng.random(shape) > 0.7).astype(np.float64)
    return rhs, cc


def _coeffs(dg, cc):
    # abcei's r1 = (rho1 - rho2) / 2, r2 = rho2 for water/air (main.f:3582-3615)
    a, b, c, ei = (np.zeros_like(cc) for _ in range(4))
    rho2 = 1.2e-3
    dg.abcei(a, b, c, ei, cc, 0.5 * (1.0 - rho2), rho2)
    return a, b, c


def test_nan_velocity_fails_the_step(dg):
    print("case 18^3 NaN in u before a step", flush=True)
    g = dg.DropletGPU(inputdata_text=inputdata_text(nx=18, ny=18, nz=18))
    try:
        g.step(0)
        u = g.get("u")
        u[9, 9, 9] = np.nan
        g.set("u", u)
        with pytest.raises(dg.DropletError, match="non-finite velocity"):
            g.step(1)
    finally:
        g.close()


@pytest.mark.parametrize("mode", [1, 2])
def test_nan_right_hand_side_fails_the_solve(dg, mode):
    print("case 18^3 NaN right-hand side, mode", mode, flush=True)
    rhs, cc = _seeded_system(18, 5)
    a, b, c = _coeffs(dg, cc)
    rhs[7, 8, 9] = np.nan
    with pytest.raises(dg.DropletError, match="non-finite"):
        dg.poisson_solve(rhs, a, b, c, mode=mode, tol=1e-8, maxit=50)


@pytest.mark.parametrize("mode", [1, 2])
def test_unconverged_solve_fails(dg, mode):
    print("case 34^3 maxit 1 at tol 1e-14, mode", mode, flush=True)
    rhs, cc = _seeded_system(34, 6)
    a, b, c = _coeffs(dg, cc)
    with pytest.raises(dg.DropletError, match="has not reached tol"):
        dg.poisson_solve(rhs, a, b, c, mode=mode, tol=1e-14, maxit=1)
    # the same system with room to converge is fine
    p, it, rel = dg.poisson_solve(rhs, a, b, c, mode=mode, tol=1e-8, maxit=200)
    assert rel < 1e-8 and it >= 1


def test_unconverged_step_fails_and_keeps_diagnostics(dg):
    print("case 18^3 step with maxit 1 at tol 1e-14", flush=True)
    g = dg.DropletGPU(inputdata_text=inputdata_text(nx=18, ny=18, nz=18))
    try:
        g.set_solver(mode=1, tol=1e-14, nu1=2, nu2=2, maxit=1, guess=0)
        with pytest.raises(dg.DropletError, match="has not reached tol"):
            g.step(0)
        st = g.state()
        assert st["mg_iters"] == 1 and st["mg_relres"] >= 1e-14
    finally:
        g.close()


def test_poisson_solve_between_steps_leaves_the_step_unchanged(dg):
    """A standalone solve on another grid between two resident steps (its
    own coarse factor, of a different size) must not disturb the resident
    solve's captured graph: the two steps match a run without it bit for bit."""
    print("case 34^3 steps with a 66x34x34 poisson_solve in between", flush=True)
    txt = inputdata_text(nx=34, ny=34, nz=34)
    rhs, cc = _seeded_system(34, 7)
    rhs66 = np.concatenate([rhs, rhs], axis=2)[:, :, :66].copy()
    cc66 = np.concatenate([cc, cc], axis=2)[:, :, :66].copy()
    a, b, c = _coeffs(dg, cc66)
    out = []
    for between in (False, True):
        g = dg.DropletGPU(inputdata_text=txt)
        g.set_vof_guard(True)
        g.set_solver(mode=2, tol=1e-8, nu1=2, nu2=2, maxit=200, guess=0)
        g.step(0)
        if between:
            p, it, rel = dg.poisson_solve(rhs66, a, b, c, mode=2, tol=1e-8, maxit=200)
            assert rel < 1e-8
        g.step(1)
        out.append({f: g.get(f) for f in ("p", "u", "v", "w", "cc")})
        g.close()
    for f in out[0]:
        assert np.array_equal(out[0][f], out[1][f]), f
