"""Multigrid pieces of the CPU restatement pinned by algebra and by an
independent sparse direct solve (scipy) of the pressure equation
    h^-2 sum_f a_f (p_nb - p) = rhs        (resid, main.f:3917-3923)
with bc_scalar ghosts (main.f:4298-4329): Neumann on every face but x-max,
where the ghost value is 0."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle_ffi import Oracle, OracleKernels as OK, inputdata_text


def assemble(a, b, c, h2i):
    """Independent assembly of the operator above as a sparse matrix."""
    nz, ny, nx = a.shape
    mx, my, mz = nx - 2, ny - 2, nz - 2
    idx = lambda i, j, k: (i - 2) + mx * ((j - 2) + my * (k - 2))  # noqa: E731
    rows, cols, vals = [], [], []
    for k in range(2, nz):
        for j in range(2, ny):
            for i in range(2, nx):
                r = idx(i, j, k)
                diag = 0.0
                for (ci, cj, ck, ii, jj, kk) in ((i, j, k, i + 1, j, k), (i - 1, j, k, i - 1, j, k),
                                                 (i, j, k, i, j + 1, k), (i, j - 1, k, i, j - 1, k),
                                                 (i, j, k, i, j, k + 1), (i, j, k - 1, i, j, k - 1)):
                    arr = a if ii != i else (b if jj != j else c)
                    coef = arr[ck - 1, cj - 1, ci - 1] * h2i
                    if ii == nx:          # Dirichlet ghost 0
                        diag -= coef
                    elif ii == 1 or jj in (1, ny) or kk in (1, nz):
                        continue          # Neumann: the face drops out
                    else:
                        diag -= coef
                        rows.append(r)
                        cols.append(idx(ii, jj, kk))
                        vals.append(coef)
                rows.append(r)
                cols.append(r)
                vals.append(diag)
    n = mx * my * mz
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def droplet_system(n=18, seed=0):
    txt = inputdata_text(nx=n, ny=n, nz=n)
    o = Oracle(txt)
    rng = np.random.default_rng(seed)
    # coefficients from the droplet's volume fraction, as abcei builds them
    a, b, c, ei = (o.field(k) for k in ("a", "b", "c", "ei"))
    OK.abcei(a, b, c, ei, o.field("cc"), 0.5 * (1.0 - o.st.rho2), o.st.rho2)
    rhs = o.field("rhs")
    rhs[...] = 0.0
    rhs[1:-1, 1:-1, 1:-1] = rng.normal(size=(n - 2, n - 2, n - 2))
    return o


def test_restrict_constant_and_interp_inverse(oracle):
    nx = ny = nz = 18
    uf = np.full((nz, ny, nx), 2.5)
    uc = np.zeros((10, 10, 10))
    OK.rstrct(uc, uf)
    assert np.all(uc[1:-1, 1:-1, 1:-1] == 2.5)
    rng = np.random.default_rng(2)
    uc2 = rng.normal(size=(10, 10, 10))
    f = np.zeros((nz, ny, nx))
    OK.interp(f, uc2)
    back = np.zeros((10, 10, 10))
    OK.rstrct(back, f)
    np.testing.assert_allclose(back[1:-1, 1:-1, 1:-1], uc2[1:-1, 1:-1, 1:-1], rtol=0, atol=1e-15)


def test_resid_of_direct_solution_is_zero(oracle):
    o = droplet_system()
    nx = o.nx
    h2i = 1.0 / (1.0 / (nx - 2)) ** 2
    a, b, c, rhs = o.field("a"), o.field("b"), o.field("c"), o.field("rhs")
    A = assemble(a, b, c, h2i)
    x = spla.spsolve(A.tocsc(), rhs[1:-1, 1:-1, 1:-1].ravel())
    p = np.zeros_like(rhs)
    p[1:-1, 1:-1, 1:-1] = x.reshape(nx - 2, nx - 2, nx - 2)
    OK.bc_scalar(p)
    res = np.zeros_like(rhs)
    OK.resid(1.0, res, p, rhs, a, b, c)
    assert np.linalg.norm(res[1:-1, 1:-1, 1:-1]) <= 1e-9 * np.linalg.norm(rhs)


def test_pcg_matches_direct_solve(oracle):
    o = droplet_system(seed=3)
    nx = o.nx
    h2i = 1.0 / (1.0 / (nx - 2)) ** 2
    a, b, c, rhs = (o.field(k).copy() for k in ("a", "b", "c", "rhs"))
    x = spla.spsolve(assemble(a, b, c, h2i).tocsc(), rhs[1:-1, 1:-1, 1:-1].ravel()).reshape(nx - 2, nx - 2, nx - 2)
    for mode in (1, 2):
        o.st.mg_mode, o.st.mg_outer, o.st.mg_prolong = mode, 1, 0
        o.st.mg_tol, o.st.mg_nu1, o.st.mg_nu2, o.st.mg_maxcycles, o.st.mg_guess = 1e-11, 2, 2, 100, 0
        it = oracle.lib().or_mg_solve_tol(o.h)
        p = o.field("p")[1:-1, 1:-1, 1:-1]
        err = np.linalg.norm(p - x) / np.linalg.norm(x)
        assert it < 40 and o.st.mg_relres_last < 1e-11
        assert err < 1e-7, (mode, err)


def test_reference_relax_converges_to_direct_solution(oracle):
    """Many reference relax sweeps (ghost = self on the Neumann faces) reach
    the same fixed point as the direct solve."""
    o = droplet_system(n=10, seed=4)
    nx = o.nx
    h2i = 1.0 / (1.0 / (nx - 2)) ** 2
    a, b, c, ei, rhs = (o.field(k).copy() for k in ("a", "b", "c", "ei", "rhs"))
    x = spla.spsolve(assemble(a, b, c, h2i).tocsc(), rhs[1:-1, 1:-1, 1:-1].ravel())
    u = np.zeros_like(rhs)
    for _ in range(20000):
        OK.relax(u, 1.0, rhs, a, b, c, ei, 1.0)
    got = u[1:-1, 1:-1, 1:-1].ravel()
    assert np.linalg.norm(got - x) / np.linalg.norm(x) < 1e-6


def _tol_solve(oracle, o, blk=(0, 0, 0), ov=0, minc=0, mode=1, tol=1e-11, nuc=2):
    o.st.mg_mode, o.st.mg_outer, o.st.mg_prolong = mode, 1, 0
    o.st.mg_tol, o.st.mg_nu1, o.st.mg_nu2, o.st.mg_maxcycles, o.st.mg_guess = tol, 2, 2, 100, 0
    o.st.mg_omega, o.st.mg_nuc = 1.15, nuc
    oracle.lib().or_set_block_smoother(o.h, *blk, minc)
    oracle.lib().or_set_block_overlap(o.h, ov)
    it = oracle.lib().or_mg_solve_tol(o.h)
    return o.field("p").copy(), it


def test_tile_smoother_with_one_tile_is_the_global_smoother(oracle):
    """A tile that covers the whole level (with or without overlap) sees no
    neighbour outside itself, so the tile smoother is the global red-black
    smoother: the same p to the bit, the same iteration count."""
    o = droplet_system(n=34, seed=6)
    p0, it0 = _tol_solve(oracle, o)
    for ov in (0, 1, 3):
        p1, it1 = _tol_solve(oracle, o, blk=(64, 64, 64), ov=ov)
        assert it1 == it0 and np.array_equal(p1, p0), ov


def test_tile_smoother_with_full_overlap_is_the_global_smoother(oracle):
    """With an overlap at least the level's extent every tile smooths the
    whole level and keeps its own cells: again the global smoother, so the
    write-back of each tile's own cells and the grown region are pinned."""
    o = droplet_system(n=34, seed=8)
    p0, it0 = _tol_solve(oracle, o)
    p1, it1 = _tol_solve(oracle, o, blk=(8, 4, 4), ov=32)
    assert it1 == it0 and np.array_equal(p1, p0)
    # a small overlap is a different (Schwarz) smoother
    p2, it2 = _tol_solve(oracle, o, blk=(8, 4, 4), ov=1)
    assert not np.array_equal(p2, p0)


def test_tile_smoother_pcg_matches_direct_solve(oracle):
    """The PCG with the tile smoother (overlap 0 and 1, FP64 and FP32
    V-cycles, tiles clipped at the level's ends) reaches the direct solution."""
    o = droplet_system(n=34, seed=9)
    nx = o.nx
    h2i = 1.0 / (1.0 / (nx - 2)) ** 2
    a, b, c, rhs = (o.field(k).copy() for k in ("a", "b", "c", "rhs"))
    x = spla.spsolve(assemble(a, b, c, h2i).tocsc(), rhs[1:-1, 1:-1, 1:-1].ravel()).reshape(nx - 2, nx - 2, nx - 2)
    for mode in (1, 2):
        for ov in (0, 1):
            p, it = _tol_solve(oracle, o, blk=(12, 6, 6), ov=ov, minc=64, mode=mode)
            err = np.linalg.norm(p[1:-1, 1:-1, 1:-1] - x) / np.linalg.norm(x)
            assert it < 40 and o.st.mg_relres_last < 1e-11, (mode, ov, it)
            assert err < 1e-7, (mode, ov, err)
