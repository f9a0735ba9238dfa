"""Tile smoother of the tolerance-mode V-cycle (dg_set_tile_smoother): the
levels between the finest and the single-workgroup tail smooth each tile
(16x8x8 cells, 32x16x8 on levels of more than 2^18 cells) on its own
(restricted additive Schwarz with an overlap of one cell, red-black SOR
inside), one launch per level visit.

The FP64 solve follows the CPU statement of the same smoother
(oracle/mg_tol.c, or_set_block_smoother(16, 8, 8, 4096),
or_set_block_large(32, 16, 8, 2^18), overlap 1):
solutions to 1e-8 and iterations within one.  The FP32-preconditioned solve
with tiles matches the FP64 solve without them to 1e-6 and returns its true
FP64 residual; graph replay, look-ahead and launch-by-launch agree to the
bit; a whole time step with tiles matches one without them to 1e-6.
"""
import numpy as np
import pytest

from oracle_ffi import Oracle, OracleKernels as OK, inputdata_text
from test_gpu_poisson import sphere_system

pytestmark = pytest.mark.gpu


def box_system(shape, seed):
    """(rhs, a, b, c) of a sphere of water in air on an nx x ny x nz grid
    with ghosts, as sphere_system builds them on a cube."""
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    h = 1.0 / (nx - 2)
    r = rng.uniform(0.2, 0.35) * (min(shape) - 2) * h
    ctr = np.array([rng.uniform(0.25, 0.75) * (m - 2) * h for m in shape])
    xs = [(np.arange(m) - 1.5) * h for m in shape]
    cc = np.zeros((nz, ny, nx))
    sub = (np.arange(4) + 0.5) / 4 * h - 0.5 * h
    for dz in sub:
        for dy in sub:
            for dx in sub:
                Z, Y, X = np.meshgrid(xs[2] + dz, xs[1] + dy, xs[0] + dx, indexing="ij")
                cc += ((X - ctr[0]) ** 2 + (Y - ctr[1]) ** 2 + (Z - ctr[2]) ** 2 < r * r)
    cc /= 64.0
    a, b, c, ei = (np.zeros((nz, ny, nx)) for _ in range(4))
    rho2 = 1.2 / 1000.0
    OK.abcei(a, b, c, ei, cc, 0.5 * (1.0 - rho2), rho2)
    vel = [rng.normal(0.0, 0.5, size=(nz, ny, nx)) for _ in range(3)]
    for q in range(3):
        for ax in range(3):
            for _ in range(2):
                vel[q] = 0.5 * (vel[q] + np.roll(vel[q], 1, axis=ax))
    u, v, w = vel
    rhs = np.zeros((nz, ny, nx))
    rhs[1:-1, 1:-1, 1:-1] = ((u[1:-1, 1:-1, 2:] - u[1:-1, 1:-1, 1:-1]) + (v[1:-1, 2:, 1:-1] - v[1:-1, 1:-1, 1:-1]) +
                             (w[2:, 1:-1, 1:-1] - w[1:-1, 1:-1, 1:-1])) / (h * h)
    return rhs, a, b, c


def oracle_solve(oracle, n, rhs, a, b, c, tiles, tol=1e-10, nuc=4, omega=1.15, shape=None, tile_nu=0):
    """The CPU statement's FP64 solve; tiles: the GPU's tile-smoother mode
    (0 none, 1 levels of 4096 < cells <= 2^18, 2 all levels above 4096)."""
    nx, ny, nz = shape or (n, n, n)
    o = Oracle(inputdata_text(nx=nx, ny=ny, nz=nz))
    for name, arr in (("rhs", rhs), ("a", a), ("b", b), ("c", c)):
        o.field(name)[...] = arr
    o.st.mg_mode, o.st.mg_outer, o.st.mg_prolong = 1, 1, 0
    o.st.mg_tol, o.st.mg_nu1, o.st.mg_nu2, o.st.mg_maxcycles, o.st.mg_guess = tol, 2, 2, 100, 0
    o.st.mg_omega, o.st.mg_nuc = omega, nuc
    L = oracle.lib()
    if tiles:
        L.or_set_block_smoother(o.h, 16, 8, 8, 4096)
        L.or_set_block_overlap(o.h, 1)
        L.or_set_block_max(o.h, 1 << 18 if tiles == 1 else 0)
        L.or_set_block_sweeps(o.h, tile_nu)
        L.or_set_block_large(o.h, 32, 16, 8, 1 << 18)
    it = L.or_mg_solve_tol(o.h)
    p = o.field("p").copy()
    o.close()
    return p, it


def rel(x, y):
    x, y = x[1:-1, 1:-1, 1:-1], y[1:-1, 1:-1, 1:-1]
    return np.linalg.norm(x - y) / np.linalg.norm(y)


@pytest.mark.parametrize("shape,tiles,tile_nu", [((66, 66, 66), 2, 0), ((130, 130, 130), 2, 0),
                                                 ((130, 130, 130), 2, 8), ((258, 130, 130), 1, 8),
                                                 ((258, 130, 130), 2, 0)])
def test_fp64_tiles_match_oracle(dg, oracle, shape, tiles, tile_nu):
    """At 256x128x128 the 128x64x64 level has more than 2^18 cells: mode 1
    keeps the global passes there, mode 2 tiles it.  tile_nu: the tile
    levels' own sweep count (0: the coarse levels' 4)."""
    print("case tile smoother vs oracle", shape, "mode", tiles, "sweeps", tile_nu, flush=True)
    rhs, a, b, c = box_system(shape, 13)
    po, ito = oracle_solve(oracle, None, rhs, a, b, c, tiles=tiles, shape=shape, tile_nu=tile_nu)
    try:
        dg.set_smoother(1.15, 4)
        dg.set_tile_smoother(tiles, tile_nu)
        pg, it, r = dg.poisson_solve(rhs, a, b, c, mode=1, tol=1e-10)
    finally:
        dg.set_tile_smoother(False)
        dg.set_smoother()
    print("iterations gpu", it, "oracle", ito, "rel", rel(pg, po), flush=True)
    assert r < 1e-10
    assert abs(it - ito) <= 1
    assert rel(pg, po) < 1e-8


def test_fp64_tiles_on_clipped_levels_match_oracle(dg, oracle):
    """96 = 3 * 2^5 interior cells: the 48^3 and 24^3 levels are not
    multiples of the tile, so tiles at their ends are clipped."""
    n = 98
    print("case tile smoother, clipped tiles", n, flush=True)
    rhs, a, b, c = sphere_system(n, 17)
    po, ito = oracle_solve(oracle, n, rhs, a, b, c, tiles=2)
    try:
        dg.set_smoother(1.15, 4)
        dg.set_tile_smoother(2)
        pg, it, r = dg.poisson_solve(rhs, a, b, c, mode=1, tol=1e-10)
    finally:
        dg.set_tile_smoother(False)
        dg.set_smoother()
    print("iterations gpu", it, "oracle", ito, "rel", rel(pg, po), flush=True)
    assert r < 1e-10 and abs(it - ito) <= 1
    assert rel(pg, po) < 1e-8


@pytest.mark.parametrize("n", [66, 130])
def test_mixed_tiles_match_fp64(dg, n):
    print("case mixed tiles vs fp64", n, flush=True)
    rhs, a, b, c = sphere_system(n, 7)
    try:
        dg.set_smoother(1.15, 4)
        p64, it64, r64 = dg.poisson_solve(rhs, a, b, c, mode=1, tol=1e-8)
        dg.set_tile_smoother(2)
        p32, it32, r32 = dg.poisson_solve(rhs, a, b, c, mode=2, tol=1e-8)
    finally:
        dg.set_tile_smoother(False)
        dg.set_smoother()
    print("iterations fp64", it64, "mixed tiles", it32, flush=True)
    assert r64 < 1e-8 and r32 < 1e-8 and it32 <= it64 + 2
    assert rel(p32, p64) < 1e-6
    res = np.zeros_like(rhs)
    OK.resid(1.0, res, p32, rhs, a, b, c)
    assert np.linalg.norm(res[1:-1, 1:-1, 1:-1]) / np.linalg.norm(rhs[1:-1, 1:-1, 1:-1]) < 2e-8


@pytest.mark.parametrize("mode", [1, 2])
def test_tiles_graph_and_lookahead_bit_identical(dg, mode):
    print("case tiles graph/look-ahead, mode", mode, flush=True)
    rhs, a, b, c = sphere_system(66, 5)
    out = []
    try:
        dg.set_smoother(1.15, 4)
        dg.set_tile_smoother(2)
        for graphs in (True, False):
            dg.set_tuning(graphs=graphs)
            for ratio in (0.0, 1e3):
                dg.set_lookahead(ratio)
                out.append(dg.poisson_solve(rhs, a, b, c, mode=mode, tol=1e-9))
    finally:
        dg.set_lookahead()
        dg.set_tuning()
        dg.set_tile_smoother(False)
        dg.set_smoother()
    p0, it0, r0 = out[0]
    assert r0 < 1e-9
    for p, it, r in out[1:]:
        assert it == it0 and r == r0 and np.array_equal(p, p0)


def test_step_with_tiles_matches_fp64(dg):
    """Three time steps of a 66^3 droplet with the mixed solver and tiles
    against the FP64 solver without them: the fields to 1e-6."""
    print("case step with tiles", flush=True)
    txt = inputdata_text(nx=66, ny=66, nz=66)
    fields = []
    try:
        dg.set_smoother(1.15, 4)
        for mode, tiles in ((1, False), (2, True)):
            dg.set_tile_smoother(tiles)
            g = dg.DropletGPU(inputdata_text=txt)
            g.clip_cc()
            g.set_vof_guard(True)
            g.set_solver(mode=mode, tol=1e-10, nu1=2, nu2=2, maxit=200, guess=0)
            for t in range(3):
                g.step(t)
                assert g.state()["mg_relres"] < 1e-10
            fields.append({f: g.get(f) for f in ("cc", "u", "v", "w", "p")})
            g.close()
    finally:
        dg.set_tile_smoother(False)
        dg.set_smoother()
    for f in ("cc", "u", "v", "w", "p"):
        e = rel(fields[1][f], fields[0][f])
        print(f, "rel", e, flush=True)
        assert e < 1e-6, (f, e)
