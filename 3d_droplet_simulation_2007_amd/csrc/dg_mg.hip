// dg_mg.hip -- pressure Poisson solvers.
//
// 1. The reference's full-multigrid schedule (mglin, main.f:3645-3763) with
//    its kernels relax/resid/rstrct/interp/addint/acfromac/slvsml, FP64,
//    same arithmetic order (parity mode).
// 2. The tolerance mode (no reference counterpart): preconditioned
//    conjugate gradients on the reference's finest-level operator, stopped
//    at ||b - A p||_2 / ||b||_2 < tol (FP64), preconditioned by one
//    symmetric V-cycle (red-black Gauss-Seidel, 8-child average restriction,
//    injection prolongation, acfromac coarse operators, FP64 Cholesky on the
//    coarsest level).  mg_mode 2 runs the V-cycle in FP32.

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <vector>

#include "dg_coarse.h"
#include "dg_ctx.h"
#include "dg_nn.h"
#include "dg_prof.h"
#include "dg_rb.h"
#include "dg_util.h"

namespace dg {

// ---------------------------------------------------------------------------
// reference kernels

// one red-black colour pass of relax (main.f:3966-3991); colour 0 updates
// cells with i+j+k even (the first pass, ksw=1)
__global__ void __launch_bounds__(256) k_relax_colour(Grid g, double *u, const double *rhs, const double *a,
                                                      const double *b, const double *c, const double *ei, double h2,
                                                      double omega, int colour)
{
    const uint3 tile = xcd_tile();
    const int j = (int)(tile.y * blockDim.y + threadIdx.y) + 2;
    const int k = (int)tile.z + 2;
    const int i = 2 * (int)(tile.x * blockDim.x + threadIdx.x) + 2 + ((colour + j + k) & 1);
    if (i > g.nx - 1 || j > g.ny - 1) return;
    const int o = g.at(i, j, k), sy = g.sy, sz = g.sz;
    double temp = (a[o] * u[o + 1] + a[o - 1] * u[o - 1] + b[o] * u[o + sy] + b[o - sy] * u[o - sy] + c[o] * u[o + sz] +
                   c[o - sz] * u[o - sz] - h2 * rhs[o]) *
                  ei[o];
    u[o] = omega * temp + (1.0 - omega) * u[o];
}

int mg_relax(const Grid &g, hipStream_t s, double *u, double lengthsize, const double *rhs, const double *a,
             const double *b, const double *c, const double *ei, double omega)
{
    const double h1 = lengthsize / (double)(g.nx - 2);
    const double h2 = h1 * h1;
    const int half = (g.nx - 2 + 1) / 2;
    dim3 grid((unsigned)((half + BX - 1) / BX), (unsigned)((g.ny - 2 + BY - 1) / BY), (unsigned)(g.nz - 2));
    for (int pass = 0; pass < 2; pass++) {
        if (op_bc_scalar(g, s, u)) return -1;
        prof_begin(PK_RELAX, s);
        hipLaunchKernelGGL(k_relax_colour, grid, interior_block(), 0, s, g, u, rhs, a, b, c, ei, h2, omega, pass);
        prof_end(PK_RELAX, s, 3.5 * 8.0 * (double)(g.nx - 2) * (g.ny - 2) * (g.nz - 2));
        DG_CHECK_LAUNCH();
    }
    return op_bc_scalar(g, s, u);
}

int mg_resid(const Grid &g, hipStream_t s, double lengthsize, double *res, const double *u, const double *rhs,
             const double *a, const double *b, const double *c)
{
    const double h1 = lengthsize / (double)(g.nx - 2);
    const double h2i = 1.0 / (h1 * h1);
    const int sy = g.sy, sz = g.sz;
    if (launch_interior(g, s, [=] __device__(int i, int j, int k) {
            const int o = g.at(i, j, k);
            const double u0 = u[o];
            res[o] = -h2i * (a[o] * (u[o + 1] - u0) + a[o - 1] * (u[o - 1] - u0) + b[o] * (u[o + sy] - u0) +
                             b[o - sy] * (u[o - sy] - u0) + c[o] * (u[o + sz] - u0) + c[o - sz] * (u[o - sz] - u0)) +
                     rhs[o];
        }))
        return -1;
    return op_bc_resid(g, s, res);
}

int mg_rstrct(const Grid &gc, hipStream_t s, double *uc, const double *uf)
{
    const Grid gf = make_grid(2 * gc.nx - 2, 2 * gc.ny - 2, 2 * gc.nz - 2);
    if (launch_interior(gc, s, [=] __device__(int ic, int jc, int kc) {
            const int f = gf.at(2 * ic - 2, 2 * jc - 2, 2 * kc - 2), sy = gf.sy, sz = gf.sz;
            const double t = uf[f] + uf[f + 1] + uf[f + sy] + uf[f + 1 + sy] + uf[f + sz] + uf[f + 1 + sz] +
                             uf[f + sy + sz] + uf[f + 1 + sy + sz];
            uc[gc.at(ic, jc, kc)] = .125 * t;
        }))
        return -1;
    return op_bc_scalar(gc, s, uc);
}

int mg_interp(const Grid &gf, hipStream_t s, double *uf, const double *uc)
{
    const Grid gc = make_grid(gf.nx / 2 + 1, gf.ny / 2 + 1, gf.nz / 2 + 1);
    if (launch_interior(gc, s, [=] __device__(int ic, int jc, int kc) {
            const double v = uc[gc.at(ic, jc, kc)];
            const int f = gf.at(2 * ic - 2, 2 * jc - 2, 2 * kc - 2), sy = gf.sy, sz = gf.sz;
            uf[f] = v; uf[f + sy] = v; uf[f + 1] = v; uf[f + 1 + sy] = v;
            uf[f + sz] = v; uf[f + sy + sz] = v; uf[f + 1 + sz] = v; uf[f + 1 + sy + sz] = v;
        }))
        return -1;
    return op_bc_scalar(gf, s, uf);
}

int mg_addint(const Grid &gf, hipStream_t s, double *uf, const double *uc, double *res)
{
    if (mg_interp(gf, s, res, uc)) return -1;
    return launch_linear(gf.n(), s, [=] __device__(size_t p) { uf[p] = uf[p] + res[p]; });
}

// acfromac (main.f:3793-3844) in one pass over the coarse interior: each
// cell forms its three face averages and, for ei, those of the faces below
// it again (the same four-term sums in the same order); the cells next to
// the lower ghost faces also write the ghost-face averages a(1,j,k),
// b(i,1,k) and c(i,j,1) that the reference's loops include.
int mg_acfromac(const Grid &gc, hipStream_t s, double *a, double *b, double *c, double *ei, const double *a2,
                const double *b2, const double *c2, void *split, int split_ct)
{
    const Grid gf = make_grid(2 * gc.nx - 2, 2 * gc.ny - 2, 2 * gc.nz - 2);
    const int sy = gc.sy, sz = gc.sz;
    const int mh = (gc.nx - 2) / 2, my = gc.ny - 2;
    const size_t half = (size_t)mh * my * (gc.nz - 2), n2 = 2 * half;
    return launch_interior(gc, s, [=] __device__(int i, int j, int k) {
        auto A = [&](int ii) {
            return 0.25 * (a2[gf.at(2 * ii - 1, 2 * j - 1, 2 * k - 1)] + a2[gf.at(2 * ii - 1, 2 * j - 2, 2 * k - 1)] +
                           a2[gf.at(2 * ii - 1, 2 * j - 1, 2 * k - 2)] + a2[gf.at(2 * ii - 1, 2 * j - 2, 2 * k - 2)]);
        };
        auto B = [&](int jj) {
            return 0.25 * (b2[gf.at(2 * i - 1, 2 * jj - 1, 2 * k - 1)] + b2[gf.at(2 * i - 2, 2 * jj - 1, 2 * k - 1)] +
                           b2[gf.at(2 * i - 1, 2 * jj - 1, 2 * k - 2)] + b2[gf.at(2 * i - 2, 2 * jj - 1, 2 * k - 2)]);
        };
        auto C = [&](int kk) {
            return 0.25 * (c2[gf.at(2 * i - 1, 2 * j - 1, 2 * kk - 1)] + c2[gf.at(2 * i - 2, 2 * j - 1, 2 * kk - 1)] +
                           c2[gf.at(2 * i - 1, 2 * j - 2, 2 * kk - 1)] + c2[gf.at(2 * i - 2, 2 * j - 2, 2 * kk - 1)]);
        };
        const int o = gc.at(i, j, k);
        const double av = A(i), aw = A(i - 1), bv = B(j), bs = B(j - 1), cv = C(k), cb = C(k - 1);
        a[o] = av;
        b[o] = bv;
        c[o] = cv;
        if (i == 2) a[o - 1] = aw;
        if (j == 2) b[o - sy] = bs;
        if (k == 2) c[o - sz] = cb;
        ei[o] = 1.0 / (av + aw + bv + bs + cv + cb);
        if (split) {  // the split V-cycle's set-up conversion (k_rb_coef) of this level
            const size_t e = (size_t)((i + j + k) & 1) * half + ((size_t)(k - 2) * my + (j - 2)) * mh + ((i - 2) >> 1);
            if (split_ct == RB_CT_F16) {
                _Float16 *p = (_Float16 *)split;
                p[e] = (_Float16)av;
                p[n2 + e] = (_Float16)bv;
                p[2 * n2 + e] = (_Float16)cv;
            } else if (split_ct == RB_CT_F64) {
                double *p = (double *)split;
                p[e] = av;
                p[n2 + e] = bv;
                p[2 * n2 + e] = cv;
            } else {
                float *p = (float *)split;
                p[e] = (float)av;
                p[n2 + e] = (float)bv;
                p[2 * n2 + e] = (float)cv;
            }
        }
    });
}

static int slvsml(Ctx &c, double lengthsize)
{
    Level &L = c.lev[0];
    for (int q = 1; q <= c.s.ntop; q++)
        if (mg_relax(L.g, c.stream, L.u, lengthsize, L.rhs, L.a, L.b, L.c, L.ei, c.s.omega)) return -1;
    return 0;
}

int op_mglin(Ctx &c)
{
    if (!c.ref_grid) {  // memoryaloc's STOP (main.f:274-322)
        set_error("STOP: nn .ne. 2**ng + 2 (the reference's multigrid schedule; use a tolerance-mode solver)");
        return -1;
    }
    hipStream_t s = c.stream;
    const int ng = c.ng;
    for (int gi = ng - 1; gi >= 1; gi--) {
        Level &C = c.lev[gi - 1], &F = c.lev[gi];
        if (mg_rstrct(C.g, s, C.rhs, F.rhs)) return -1;
        if (mg_acfromac(C.g, s, C.a, C.b, C.c, C.ei, F.a, F.b, F.c)) return -1;
    }
    Level &L1 = c.lev[0];
    if (mg_fill(L1.g.n(), s, L1.u, 0.0)) return -1;
    // the first coarsest solve is given lengthscale, not lengthsize (main.f:3689)
    if (slvsml(c, c.s.lengthscale)) return -1;
    const double ls = c.s.lengthsize;
    for (int j = 2; j <= ng; j++) {
        Level &J = c.lev[j - 1];
        if (mg_interp(J.g, s, J.u, c.lev[j - 2].u)) return -1;
        for (int jc = 1; jc <= c.s.cycles; jc++) {
            for (int jj = j; jj >= 2; jj--) {
                Level &F = c.lev[jj - 1], &C = c.lev[jj - 2];
                for (int q = 1; q <= c.s.npre; q++)
                    if (mg_relax(F.g, s, F.u, ls, F.rhs, F.a, F.b, F.c, F.ei, c.s.omega)) return -1;
                if (mg_resid(F.g, s, ls, F.res, F.u, F.rhs, F.a, F.b, F.c)) return -1;
                if (mg_rstrct(C.g, s, C.rhs, F.res)) return -1;
                if (mg_fill(C.g.n(), s, C.u, 0.0)) return -1;
            }
            if (slvsml(c, ls)) return -1;
            for (int jj = 2; jj <= j; jj++) {
                Level &F = c.lev[jj - 1], &C = c.lev[jj - 2];
                if (mg_addint(F.g, s, F.u, C.u, F.res)) return -1;
                for (int q = 1; q <= c.s.npost; q++)
                    if (mg_relax(F.g, s, F.u, ls, F.rhs, F.a, F.b, F.c, F.ei, c.s.omega)) return -1;
            }
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// tolerance mode

// bc_scalar ghost folding for the operator: across x-min, y and z faces the
// neighbour equals the cell itself (the face drops out), across x-max the
// ghost is 0.

// Gauss-Seidel colour pass with the operator's own diagonal, in place, with
// SOR weight om: u = om * gs + (1 - om) * u (om = 1: Gauss-Seidel).  From a
// zero iterate the first pass reads no neighbour (zero_init) and the first
// two passes no old value of their own cells (zero_own).  TOP only
// separates the finest level's launches into their own symbol.
template <typename T, bool TOP>
__global__ void __launch_bounds__(256) k_gs_colour(Grid g, T *u, const T *f, const T *a, const T *b, const T *c, T h2,
                                                   int colour, int zero_init, T om, int zero_own)
{
    const uint3 tile = xcd_tile();
    const int j = (int)(tile.y * blockDim.y + threadIdx.y) + 2;
    const int k = (int)tile.z + 2;
    const int i = 2 * (int)(tile.x * blockDim.x + threadIdx.x) + 2 + ((colour + j + k) & 1);
    if (i > g.nx - 1 || j > g.ny - 1) return;
    const int o = g.at(i, j, k), sy = g.sy, sz = g.sz;
    const T ae = a[o], aw = a[o - 1], bn = b[o], bs = b[o - sy], ct = c[o], cb = c[o - sz];
    T s = (T)0, d = ae;
    if (!zero_init) {
        if (i < g.nx - 1) s += ae * u[o + 1];
        if (i > 2) s += aw * u[o - 1];
        if (j < g.ny - 1) s += bn * u[o + sy];
        if (j > 2) s += bs * u[o - sy];
        if (k < g.nz - 1) s += ct * u[o + sz];
        if (k > 2) s += cb * u[o - sz];
    }
    if (i > 2) d += aw;
    if (j < g.ny - 1) d += bn;
    if (j > 2) d += bs;
    if (k < g.nz - 1) d += ct;
    if (k > 2) d += cb;
    const T gsv = (s - h2 * f[o]) / d;
    const T uold = (zero_own || om == (T)1) ? (T)0 : u[o];  // Gauss-Seidel needs no old value
    u[o] = om * gsv + ((T)1 - om) * uold;
}

// r = f - A u
template <typename T>
__global__ void __launch_bounds__(256) k_resid_fold(Grid g, T *r, const T *u, const T *f, const T *a, const T *b,
                                                    const T *c, T h2i)
{
    DG_INTERIOR_IJK(g);
    const int o = g.at(i, j, k), sy = g.sy, sz = g.sz;
    const T u0 = u[o];
    const T ue = i < g.nx - 1 ? u[o + 1] : (T)0;
    const T uw = i > 2 ? u[o - 1] : u0;
    const T un = j < g.ny - 1 ? u[o + sy] : u0;
    const T us = j > 2 ? u[o - sy] : u0;
    const T ut = k < g.nz - 1 ? u[o + sz] : u0;
    const T ub = k > 2 ? u[o - sz] : u0;
    const T s = a[o] * (ue - u0) + a[o - 1] * (uw - u0) + b[o] * (un - u0) + b[o - sy] * (us - u0) + c[o] * (ut - u0) +
                c[o - sz] * (ub - u0);
    r[o] = f[o] - h2i * s;
}

template <typename T>
__global__ void __launch_bounds__(256) k_restrict_avg(Grid gc, Grid gf, T *fc, const T *rf)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 2;
    const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) + 2;
    const int k = (int)blockIdx.z + 2;
    if (i > gc.nx - 1 || j > gc.ny - 1) return;
    const int f = gf.at(2 * i - 2, 2 * j - 2, 2 * k - 2), sy = gf.sy, sz = gf.sz;
    const T t = rf[f] + rf[f + 1] + rf[f + sy] + rf[f + 1 + sy] + rf[f + sz] + rf[f + 1 + sz] + rf[f + sy + sz] +
                rf[f + 1 + sy + sz];
    fc[gc.at(i, j, k)] = (T)0.125 * t;
}

template <typename T>
__global__ void __launch_bounds__(256) k_prolong_inj_add(Grid gf, Grid gc, T *uf, const T *uc)
{
    DG_INTERIOR_IJK(gf);
    uf[gf.at(i, j, k)] += uc[gc.at(i / 2 + 1, j / 2 + 1, k / 2 + 1)];
}

template <typename T>
__global__ void k_zero(size_t n, T *a)
{
    size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p < n) a[p] = (T)0;
}

__global__ void k_to_f32(size_t n, const double *src, float *dst)
{
    size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p < n) dst[p] = (float)src[p];
}

// coarsest level: assemble -A (SPD) densely and factor it (Cholesky), one
// block; the factor is kept for the forward/back substitutions of every
// V-cycle of the solve
// info: 0 on success, else the 1-based index of the first non-positive
// pivot (written by thread 0 only; cleared here, so every solve reports its
// own factorisation)
__global__ void __launch_bounds__(256) k_coarse_factor(Grid g, const double *a, const double *b, const double *c,
                                                      double h2i, double *M, double *info)
{
    const int mx = g.nx - 2, my = g.ny - 2, mz = g.nz - 2, n = mx * my * mz;
    if (threadIdx.x == 0) *info = 0.0;
    for (int p = threadIdx.x; p < n * n; p += blockDim.x) M[p] = 0.0;
    __syncthreads();
    for (int row = threadIdx.x; row < n; row += blockDim.x) {
        const int i = row % mx + 2, j = (row / mx) % my + 2, k = row / (mx * my) + 2;
        const int o = g.at(i, j, k);
        double d = a[o];  // x-max ghost (0) or real neighbour: the face always counts
        if (i < g.nx - 1) M[row * n + row + 1] -= a[o] * h2i;
        if (i > 2) { d += a[o - 1]; M[row * n + row - 1] -= a[o - 1] * h2i; }
        if (j < g.ny - 1) { d += b[o]; M[row * n + row + mx] -= b[o] * h2i; }
        if (j > 2) { d += b[o - g.sy]; M[row * n + row - mx] -= b[o - g.sy] * h2i; }
        if (k < g.nz - 1) { d += c[o]; M[row * n + row + mx * my] -= c[o] * h2i; }
        if (k > 2) { d += c[o - g.sz]; M[row * n + row - mx * my] -= c[o - g.sz] * h2i; }
        M[row * n + row] += d * h2i;
    }
    __syncthreads();
    // right-looking Cholesky, lower triangle
    for (int p = 0; p < n; p++) {
        if (threadIdx.x == 0) {
            double dpp = M[p * n + p];
            if (!(dpp > 0.0) && *info == 0.0) *info = (double)(p + 1);
            M[p * n + p] = sqrt(dpp);
        }
        __syncthreads();
        const double l = M[p * n + p];
        for (int r = p + 1 + threadIdx.x; r < n; r += blockDim.x) M[r * n + p] /= l;
        __syncthreads();
        for (int q = threadIdx.x; q < (n - p - 1) * (n - p - 1); q += blockDim.x) {
            int r = p + 1 + q / (n - p - 1), cc = p + 1 + q % (n - p - 1);
            if (cc <= r) M[r * n + cc] -= M[r * n + p] * M[cc * n + p];
        }
        __syncthreads();
    }
}

template <typename T>
__global__ void __launch_bounds__(64) k_coarse_solve(Grid g, const double *M, const T *f, T *u)
{
    // one wave: substitutions with the rhs staged in LDS
    __shared__ double y[512];
    coarse_subst<T>(g, M, f, u, y, (int)threadIdx.x, 64);
}

// The coarse factor belongs to the context: a captured PCG graph bakes its
// address into the coarse-solve kernels, so it lives and dies with the
// context's graphs and is never moved while they exist.
const double *coarse_factor(const Ctx &c) { return c.coarse_M; }

// coarsest-level FP32 solve (natural layout f32 -> u32) for the split V-cycle
int coarse_solve_f32(Ctx &c)
{
    Level &L = c.lev[0];
    hipLaunchKernelGGL(k_coarse_solve<float>, dim3(1), dim3(64), 0, c.stream, L.g, c.coarse_M, L.f32, L.u32);
    DG_CHECK_LAUNCH();
    return 0;
}

// the same in FP64 (f64 -> u64) for the split FP64 V-cycle
int coarse_solve_f64(Ctx &c)
{
    Level &L = c.lev[0];
    hipLaunchKernelGGL(k_coarse_solve<double>, dim3(1), dim3(64), 0, c.stream, L.g, c.coarse_M, L.f64, L.u64);
    DG_CHECK_LAUNCH();
    return 0;
}

static inline double lev_h2i(const Ctx &c, const Grid &g)
{
    const double h = c.s.lengthsize / (double)(g.nx - 2);
    return 1.0 / (h * h);
}

template <typename T>
static void level_ptrs(Level &L, T *&u, T *&f, T *&r, const T *&a, const T *&b, const T *&cc);
template <>
void level_ptrs<double>(Level &L, double *&u, double *&f, double *&r, const double *&a, const double *&b,
                        const double *&cc)
{
    u = L.u64; f = L.f64; r = L.r64; a = L.a; b = L.b; cc = L.c;
}
template <>
void level_ptrs<float>(Level &L, float *&u, float *&f, float *&r, const float *&a, const float *&b, const float *&cc)
{
    u = L.u32; f = L.f32; r = L.r32; a = L.a32; b = L.b32; cc = L.c32;
}

template <typename T>
static int gs_sweep(Ctx &c, const Grid &g, T *u, const T *f, const T *a, const T *b, const T *cc, T h2, int first,
                    int second, int zero_first, bool top, T om)
{
    const int half = (g.nx - 2 + 1) / 2;
    dim3 grid((unsigned)((half + BX - 1) / BX), (unsigned)((g.ny - 2 + BY - 1) / BY), (unsigned)(g.nz - 2));
    // algorithmic bytes: one red-black sweep moves 7 words per cell (relax's
    // a,b,c,ei,rhs,u -> u); one colour pass is half a sweep
    const double bytes = 3.5 * (double)sizeof(T) * (double)(g.nx - 2) * (g.ny - 2) * (g.nz - 2);
    if (!top) {
        hipLaunchKernelGGL((k_gs_colour<T, false>), grid, interior_block(), 0, c.stream, g, u, f, a, b, cc, h2, first,
                           zero_first, om, zero_first);
        DG_CHECK_LAUNCH();
        hipLaunchKernelGGL((k_gs_colour<T, false>), grid, interior_block(), 0, c.stream, g, u, f, a, b, cc, h2, second,
                           0, om, zero_first);
        DG_CHECK_LAUNCH();
        return 0;
    }
    // the finest level: the profiled kernel family PK_GS_COLOUR
    prof_begin(PK_GS_COLOUR, c.stream);
    hipLaunchKernelGGL((k_gs_colour<T, true>), grid, interior_block(), 0, c.stream, g, u, f, a, b, cc, h2, first,
                       zero_first, om, zero_first);
    prof_end(PK_GS_COLOUR, c.stream, bytes);
    DG_CHECK_LAUNCH();
    prof_begin(PK_GS_COLOUR, c.stream);
    hipLaunchKernelGGL((k_gs_colour<T, true>), grid, interior_block(), 0, c.stream, g, u, f, a, b, cc, h2, second, 0,
                       om, zero_first);
    prof_end(PK_GS_COLOUR, c.stream, bytes);
    DG_CHECK_LAUNCH();
    return 0;
}

// symmetric V-cycle from a zero guess on level l: u_l ~ A_l^-1 f_l
template <typename T>
static int vcycle(Ctx &c, int l)
{
    Level &L = c.lev[l];
    const Grid g = L.g;
    T *u, *f, *r;
    const T *a, *b, *cc;
    level_ptrs<T>(L, u, f, r, a, b, cc);
    if (l == 0) {
        hipLaunchKernelGGL(k_coarse_solve<T>, dim3(1), dim3(64), 0, c.stream, g, c.coarse_M, f, u);
        DG_CHECK_LAUNCH();
        return 0;
    }
    const double h2i = lev_h2i(c, g);
    const T h2 = (T)(1.0 / h2i), th2i = (T)h2i;
    const bool top = l == c.ng - 1;
    const int nu1 = (top || c.mg_nuc <= 0) ? c.mg_nu1 : c.mg_nuc;
    const int nu2 = (top || c.mg_nuc <= 0) ? c.mg_nu2 : c.mg_nuc;
    const T om = (T)c.mg_omega;
    for (int sidx = 0; sidx < nu1; sidx++)
        if (gs_sweep<T>(c, g, u, f, a, b, cc, h2, 0, 1, sidx == 0, top, om)) return -1;
    if (nu1 == 0) {
        hipLaunchKernelGGL(k_zero<T>, dim3((unsigned)((g.n() + 255) / 256)), dim3(256), 0, c.stream, g.n(), u);
        DG_CHECK_LAUNCH();
    }
    prof_begin(PK_RESID, c.stream);
    hipLaunchKernelGGL(k_resid_fold<T>, interior_grid(g), interior_block(), 0, c.stream, g, r, u, f, a, b, cc, th2i);
    prof_end(PK_RESID, c.stream, 6.0 * (double)sizeof(T) * (double)(g.nx - 2) * (g.ny - 2) * (g.nz - 2));
    DG_CHECK_LAUNCH();
    Level &C = c.lev[l - 1];
    T *uc, *fc, *rc;
    const T *ac, *bc, *ccc;
    level_ptrs<T>(C, uc, fc, rc, ac, bc, ccc);
    hipLaunchKernelGGL(k_restrict_avg<T>, interior_grid(C.g), interior_block(), 0, c.stream, C.g, g, fc, r);
    DG_CHECK_LAUNCH();
    if (vcycle<T>(c, l - 1)) return -1;
    hipLaunchKernelGGL(k_prolong_inj_add<T>, interior_grid(g), interior_block(), 0, c.stream, g, C.g, u, uc);
    DG_CHECK_LAUNCH();
    for (int sidx = 0; sidx < nu2; sidx++)
        if (gs_sweep<T>(c, g, u, f, a, b, cc, h2, 1, 0, 0, top, om)) return -1;
    return 0;
}

// ---- PCG vector kernels on the finest level (FP64) ----

// K1: partial sums of r.z and z.q (z = the V-cycle result, read through an
// accessor: natural layout of type T, or the red-black split FP32 layout)
template <class Z>
__global__ void __launch_bounds__(256) k_dot_rz_zq(Grid g, const double *r, Z z, const double *q, double *partial)
{
    __shared__ double lds4[4];
    // block b takes the interior rows b, b + gridDim.x, ...; its threads
    // stride along each row (fixed assignment: deterministic sums)
    const int my = g.ny - 2, rows = my * (g.nz - 2);
    const int G = (int)gridDim.x;
    double a1 = 0.0, a2 = 0.0;
    int rr = (int)blockIdx.x;
    if (g.nx - 2 <= 256) {
        // at most one cell of a row per thread: four rows' loads in flight
        // before their additions, which keep the row order
        const int i = 2 + (int)threadIdx.x;
        const bool act = i <= g.nx - 1;
        for (; rr + 3 * G < rows; rr += 4 * G) {
            double zv[4], rv[4], qv[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int ru = rr + u * G;
                const int j = ru % my + 2, k = ru / my + 2;
                const int o = g.at(act ? i : 2, j, k);
                zv[u] = z(act ? i : 2, j, k, o);
                rv[u] = r[o];
                qv[u] = q[o];
            }
            if (act) {
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    a1 += rv[u] * zv[u];
                    a2 += zv[u] * qv[u];
                }
            }
        }
    }
    for (; rr < rows; rr += G) {
        const int j = rr % my + 2, k = rr / my + 2;
        for (int i = 2 + (int)threadIdx.x; i <= g.nx - 1; i += 256) {
            const int o = g.at(i, j, k);
            const double zz = z(i, j, k, o);
            a1 += r[o] * zz;
            a2 += zz * q[o];
        }
    }
    double s1 = block_reduce256<false>(a1, lds4);
    double s2 = block_reduce256<false>(a2, lds4);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = s1;
        partial[2 * blockIdx.x + 1] = s2;
    }
}

__global__ void __launch_bounds__(256) k_final2(const double *partial, int nb, double *out)
{
    __shared__ double lds4[4];
    double a1 = 0.0, a2 = 0.0;
    for (int p = threadIdx.x; p < nb; p += 256) {
        a1 += partial[2 * p];
        a2 += partial[2 * p + 1];
    }
    double s1 = block_reduce256<false>(a1, lds4);
    double s2 = block_reduce256<false>(a2, lds4);
    if (threadIdx.x == 0) {
        out[0] = s1;
        out[1] = s2;
    }
}

// device scalar slots
enum { S_RZ = 16, S_ZQ, S_RZOLD, S_ALPHA, S_PQ, S_RR, S_BB, S_FIRST, S_IT, S_STOP, S_TOL, S_CINFO, S_AH = 32 };

// The search directions of a solve go to a ring of PCG_HIST buffers and
// their step lengths to ds[S_AH ..]; x takes x + alpha_j p_j for the
// buffered j in iteration order when the ring is full and at the end of
// the solve (k_pcg_flush), the same additions in the same order as one
// x += alpha p per iteration, without reading and writing x every
// iteration.  ds[S_IT]: the solve's iteration count, read by the kernels.
constexpr int PCG_HIST = 16;
static_assert(S_AH + PCG_HIST <= 64, "device scalar slots");

// K2: p_new = z + beta p_old (beta from device scalars), q = A p_new,
// partial sums of p_new . q
// planes per workgroup of k_pcg_dir, and rows of its column tile
constexpr int PCG_KCH = 8, PCG_TY = 4;

// Barrier for LDS hand-offs only: waits for this wave's LDS operations,
// not for its global loads in flight (which __syncthreads' workgroup fence
// would drain), so loads issued before it overlap the wait.
__device__ inline void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// A workgroup of 64 x PCG_TY threads owns a 64 x PCG_TY column tile and marches up
// PCG_KCH planes.  Per plane it forms p_new = z + beta p_old once for the
// tile and a one-cell ring around it in LDS, so the in-plane neighbours
// come from LDS; each thread carries its column's plane-below value and
// c-coefficient in registers.  Three plane buffers: plane k + 1 is stored
// while plane k - 1's buffer is free, so one barrier per plane separates
// the stores from the reads; the loads of plane k + 2's z and p_old and of
// plane k's coefficients are issued before that barrier and land while it
// waits.  The per-cell arithmetic is the stencil of resid (main.f:3917-3923)
// on p_new with bc_scalar folding.  PT: the storage type of the search
// direction; p_new is rounded to it before the stencil, so q is A applied
// to the stored direction and x, r stay consistent.
template <class Z, typename PT>
__global__ void __launch_bounds__(64 * PCG_TY) k_pcg_dir(Grid g, Z z, PT *hist, size_t nslot, double *q, const double *a,
                                                 const double *b, const double *c, double h2i, const double *ds,
                                                 double *partial)
{
    const int it = (int)ds[S_IT];
    PT *pnew = hist + (size_t)(it % PCG_HIST) * nslot;
    const PT *pold = hist + (size_t)((it + PCG_HIST - 1) % PCG_HIST) * nslot;
    constexpr int NT = 64 * PCG_TY, TW = 66, TH = PCG_TY + 2, NE = TH * TW, NFE = (NE + NT - 1) / NT;  // tile + ring
    __shared__ double pl[3][TH][TW];
    __shared__ double ldsw[PCG_TY];
    unsigned lin;
    const uint3 tile = xcd_tile(&lin);
    const int bidx = (int)lin;
    const int tx = (int)threadIdx.x, ty = (int)threadIdx.y, tid = tx + 64 * ty;
    const int i0 = (int)tile.x * 64 + 2, j0 = (int)tile.y * PCG_TY + 2;
    const int i = i0 + tx, j = j0 + ty;
    const int k0 = (int)tile.z * PCG_KCH + 2;
    const int k1 = min(k0 + PCG_KCH, g.nz);  // interior planes k0 .. k1 - 1 (the last is nz - 1)
    const int kl = min(k1, g.nz - 1);        // the last plane whose p_new is formed
    const bool first = ds[S_FIRST] != 0.0;
    const double beta = first ? 0.0 : -ds[S_ALPHA] * ds[S_ZQ] / ds[S_RZOLD];
    const int sy = g.sy, sz = g.sz;
    // the first iteration's direction is z itself (p_old is not read)
    auto pdir = [&](double zz, PT po) { return (double)(PT)(first ? zz : zz + beta * (double)po); };
    // this thread's ring places e = tid + NT u: loads of plane kk (places
    // outside the interior load the cell (2, 2, kk) and store 0: no
    // stencil reads them) and their stores into buffer buf
    int ei[NFE], ej[NFE];
    bool eok[NFE];
#pragma unroll
    for (int u = 0; u < NFE; u++) {
        const int e = tid + NT * u, yy = e / TW, xx = e - yy * TW;
        const int ii = i0 - 1 + xx, jj = j0 - 1 + yy;
        eok[u] = e < NE && ii >= 2 && ii <= g.nx - 1 && jj >= 2 && jj <= g.ny - 1;
        ei[u] = eok[u] ? ii : 2;
        ej[u] = eok[u] ? jj : 2;
    }
    double zr[NFE];
    PT pr[NFE];
    auto load = [&](int kk) {
#pragma unroll
        for (int u = 0; u < NFE; u++) {
            const int oo = g.at(ei[u], ej[u], kk);
            zr[u] = z(ei[u], ej[u], kk, oo);
            pr[u] = pold[oo];  // a ring slot in every iteration; unused in the first
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int u = 0; u < NFE; u++) {
            const int e = tid + NT * u;
            if (e < NE) (&pl[buf][0][0])[e] = eok[u] ? pdir(zr[u], pr[u]) : 0.0;
        }
    };
    const bool in = i <= g.nx - 1 && j <= g.ny - 1;
    double contrib = 0.0, p0 = 0.0, pb = 0.0, cb = 0.0;
    load(k0);
    store(0);
    load(min(k0 + 1, kl));
    if (in) {
        const int o = g.at(i, j, k0);
        if (k0 > 2) {
            const int ob = o - sz;
            pb = pdir(z(i, j, k0 - 1, ob), pold[ob]);
        }
        cb = c[o - sz];
    }
    lds_barrier();
    if (in) {
        p0 = pl[0][ty + 1][tx + 1];
        if (k0 == 2) pb = p0;
    }
    int cur = 0;
    for (int k = k0; k < k1; k++) {
        const int nxt = cur == 2 ? 0 : cur + 1;
        const int o = g.at(in ? i : 2, in ? j : 2, k);
        const double ae = a[o], aw = a[o - 1], bn = b[o], bs = b[o - sy], ct = c[o];
        if (k + 1 <= kl) store(nxt);
        load(min(k + 2, kl));  // the last iterations reload plane kl: no branch around the loads
        lds_barrier();
        if (in) {
            const double pt = k < g.nz - 1 ? pl[nxt][ty + 1][tx + 1] : p0;
            const double pe = i < g.nx - 1 ? pl[cur][ty + 1][tx + 2] : 0.0;
            const double pw = i > 2 ? pl[cur][ty + 1][tx] : p0;
            const double pn = j < g.ny - 1 ? pl[cur][ty + 2][tx + 1] : p0;
            const double ps = j > 2 ? pl[cur][ty][tx + 1] : p0;
            const double s = ae * (pe - p0) + aw * (pw - p0) + bn * (pn - p0) + bs * (ps - p0) + ct * (pt - p0) +
                             cb * (pb - p0);
            const double qq = h2i * s;
            pnew[o] = (PT)p0;
            q[o] = qq;
            contrib += p0 * qq;
            pb = p0;
            p0 = pt;
            cb = ct;
        }
        cur = nxt;
    }
    // block of 64 x PCG_TY: reduce
    double v = wave_sum(contrib);
    if ((tid & 63) == 0) ldsw[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
        double t = ldsw[0];
        for (int w = 1; w < PCG_TY; w++) t += ldsw[w];
        partial[bidx] = t;
    }
}

// K3: alpha = rz/pq; r -= alpha q (x takes alpha p in k_pcg_flush); partial sums of r.r
// (fs != null: also the red-black split copy of the new r in the V-cycle's
// type PT, the next V-cycle's right-hand side)
template <typename PT>
__global__ void __launch_bounds__(256) k_pcg_update(Grid g, double *r, const double *q, double *ds, double *partial,
                                                    PT *fs, RbGeom rg)
{
    __shared__ double lds4[4];
    unsigned lin;
    const uint3 tile = xcd_tile(&lin);
    const int bidx = (int)lin;
    const int i = (int)(tile.x * blockDim.x + threadIdx.x) + 2;
    const int j = (int)(tile.y * blockDim.y + threadIdx.y) + 2;
    const int k = (int)tile.z + 2;
    const int tid = (int)(threadIdx.x + blockDim.x * threadIdx.y);
    const double alpha = ds[S_RZ] / ds[S_PQ];
    double contrib = 0.0;
    if (i <= g.nx - 1 && j <= g.ny - 1) {
        const int o = g.at(i, j, k);
        const double rn = r[o] - alpha * q[o];
        r[o] = rn;
        if (fs) fs[rg.at32(i, j, k)] = (PT)rn;
        contrib = rn * rn;
    }
    double v = wave_sum(contrib);
    if ((tid & 63) == 0) lds4[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) partial[bidx] = lds4[0] + lds4[1] + lds4[2] + lds4[3];
}

__global__ void k_final_sum(const double *partial, int nb, double *out)
{
    __shared__ double lds4[4];
    double a1 = 0.0;
    for (int p = threadIdx.x; p < nb; p += 256) a1 += partial[p];
    double s1 = block_reduce256<false>(a1, lds4);
    if (threadIdx.x == 0) *out = s1;
}

// first stage of a deterministic two-stage sum of `stride`-interleaved
// partials: block b sums entries [b*RP_CHUNK, (b+1)*RP_CHUNK) of each lane
constexpr int RP_CHUNK = 2048;
__global__ void __launch_bounds__(256) k_partial_chunks(const double *partial, int nb, int stride, double *out)
{
    __shared__ double lds4[4];
    const int lo = (int)blockIdx.x * RP_CHUNK, hi = min(lo + RP_CHUNK, nb);
    for (int comp = 0; comp < stride; comp++) {
        double a1 = 0.0;
        for (int p = lo + (int)threadIdx.x; p < hi; p += 256) a1 += partial[stride * p + comp];
        double s1 = block_reduce256<false>(a1, lds4);
        if (threadIdx.x == 0) out[stride * blockIdx.x + comp] = s1;
    }
}

// sums nb partials (stride 1) or nb pairs (stride 2) into out[0..stride)
static int sum_partials(hipStream_t s, const double *partial, int nb, int stride, double *scratch, double *out)
{
    const int nchunk = (nb + RP_CHUNK - 1) / RP_CHUNK;
    hipLaunchKernelGGL(k_partial_chunks, dim3((unsigned)nchunk), dim3(256), 0, s, partial, nb, stride, scratch);
    DG_CHECK_LAUNCH();
    if (stride == 2)
        hipLaunchKernelGGL(k_final2, dim3(1), dim3(256), 0, s, scratch, nchunk, out);
    else
        hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(256), 0, s, scratch, nchunk, out);
    DG_CHECK_LAUNCH();
    return 0;
}

// scalar bookkeeping after K3: alpha_old = rz/pq, rz_old = rz, first = 0,
// and the stopping test on the same doubles as the host's: stop = 1 when
// |r| / |b| < tol, 2 when r.r is not finite (the solve has failed; the host
// reports it), 0 otherwise.  An iteration that finds stop set was queued ahead of
// the host's test and comes after convergence: it records nothing, so the
// flushes and the returned residual see only the iterations before it.
// poll: mapped host memory; r.r and the stopping flag into the pair of
// slots of the iteration number's parity (the host is at most one
// iteration behind the GPU), then, after a system-scope fence, the
// iteration count (the host waits for the count, then reads the pair)
__device__ inline void pcg_scalars(double *ds, double *poll)
{
    if (ds[S_STOP] != 0.0) return;
    double bnorm = sqrt(ds[S_BB]);
    if (bnorm == 0.0) bnorm = 1.0;
    const double rr = ds[S_RR];
    const double stop = !isfinite(rr) ? 2.0 : (sqrt(rr) / bnorm < ds[S_TOL] ? 1.0 : 0.0);
    ds[S_STOP] = stop;
    const int it = (int)ds[S_IT];
    double *pair = poll + 2 + 2 * ((it + 1) & 1);
    pair[0] = ds[S_RR];
    pair[1] = stop;
    ds[S_ALPHA] = ds[S_RZ] / ds[S_PQ];
    ds[S_AH + it % PCG_HIST] = ds[S_ALPHA];
    ds[S_IT] = (double)(it + 1);
    ds[S_RZOLD] = ds[S_RZ];
    ds[S_FIRST] = 0.0;
    __threadfence_system();
    poll[1] = (double)(it + 1);
}

// the second stage of r.r (its chunk sums, in order) and the scalar
// bookkeeping in one launch
__global__ void __launch_bounds__(256) k_final_scalars(const double *chunks, int nchunk, double *ds, double *poll)
{
    __shared__ double lds4[4];
    double a1 = 0.0;
    for (int p = threadIdx.x; p < nchunk; p += 256) a1 += chunks[p];
    const double s1 = block_reduce256<false>(a1, lds4);
    if (threadIdx.x == 0) {
        ds[S_RR] = s1;
        pcg_scalars(ds, poll);
    }
}

// a one-launch deterministic sum of nb partials by one workgroup of 1024
// threads (thread t adds entries t, t + 1024, ...; the 16 wave sums are
// added in order)
__global__ void __launch_bounds__(1024) k_sum_wide(const double *partial, int nb, double *out)
{
    __shared__ double lds16[16];
    double a1 = 0.0;
    for (int p = threadIdx.x; p < nb; p += 1024) a1 += partial[p];
    a1 = wave_sum(a1);
    if ((threadIdx.x & 63) == 0) lds16[threadIdx.x >> 6] = a1;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = lds16[0];
        for (int q = 1; q < 16; q++) r += lds16[q];
        *out = r;
    }
}

// x += alpha_j p_j for the first N buffers of the ring, in order.  N is a
// template parameter so that a cell's N direction loads are all in flight
// before the first addition (a run-time count left one load per addition)
template <typename PT, int N>
__global__ void __launch_bounds__(256) k_pcg_flush(Grid g, double *x, const PT *hist, size_t nslot, const double *ds)
{
    DG_INTERIOR_IJK(g);
    const int o = g.at(i, j, k);
    PT h[N];
#pragma unroll
    for (int q = 0; q < N; q++) h[q] = hist[(size_t)q * nslot + o];
    double v = x[o];
#pragma unroll
    for (int q = 0; q < N; q++) v = v + ds[S_AH + q] * (double)h[q];
    x[o] = v;
}

template <typename PT>
static int launch_flush(int count, hipStream_t s, const Grid &g, double *x, const PT *hist, size_t nslot,
                        const double *ds)
{
    static_assert(PCG_HIST == 16, "one instantiation per ring fill");
    const dim3 ig = interior_grid(g), ib = interior_block();
    switch (count) {
#define DG_FLUSH(N)                                                                                   \
    case N:                                                                                           \
        hipLaunchKernelGGL((k_pcg_flush<PT, N>), ig, ib, 0, s, g, x, hist, nslot, ds);                \
        break;
        DG_FLUSH(1) DG_FLUSH(2) DG_FLUSH(3) DG_FLUSH(4) DG_FLUSH(5) DG_FLUSH(6) DG_FLUSH(7) DG_FLUSH(8)
        DG_FLUSH(9) DG_FLUSH(10) DG_FLUSH(11) DG_FLUSH(12) DG_FLUSH(13) DG_FLUSH(14) DG_FLUSH(15) DG_FLUSH(16)
#undef DG_FLUSH
    default:
        return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// the previous-step guess is dropped for the zero guess when its residual
// is no smaller than b (r.r >= b.b): x = 0, r = b
__global__ void __launch_bounds__(256) k_guess_check(Grid g, double *x, double *r, const double *bvec,
                                                     const double *ds)
{
    if (ds[S_RR] < ds[S_BB]) return;
    DG_INTERIOR_IJK(g);
    const int o = g.at(i, j, k);
    x[o] = 0.0;
    r[o] = bvec[o];
}

__global__ void k_guess_check_scalar(double *ds)
{
    if (ds[S_RR] >= ds[S_BB]) ds[S_RR] = ds[S_BB];
}

// r = b - A x with folded ghosts; partial sums of r.r and b.b
// ZERO: x is zero, so A x is +0 on every cell and r = b - h2i * 0 = b bit
// for bit; x and the coefficients are not read, and fs (when given) takes
// the split copy of r in the V-cycle's type FT, its first right-hand side
template <bool ZERO, typename FT>
__global__ void __launch_bounds__(256) k_init_resid(Grid g, double *r, const double *x, const double *bvec,
                                                    const double *a, const double *b, const double *c, double h2i,
                                                    double *partial, FT *fs, RbGeom rg)
{
    __shared__ double lds4[4];
    __shared__ double lds4b[4];
    unsigned lin;
    const uint3 tile = xcd_tile(&lin);
    const int bidx = (int)lin;
    const int i = (int)(tile.x * blockDim.x + threadIdx.x) + 2;
    const int j = (int)(tile.y * blockDim.y + threadIdx.y) + 2;
    const int k = (int)tile.z + 2;
    const int tid = (int)(threadIdx.x + blockDim.x * threadIdx.y);
    double c1 = 0.0, c2 = 0.0;
    if (ZERO && i <= g.nx - 1 && j <= g.ny - 1) {
        const double bb = bvec[g.at(i, j, k)];
        r[g.at(i, j, k)] = bb;
        if (fs) fs[rg.at32(i, j, k)] = (FT)bb;
        c1 = bb * bb;
        c2 = bb * bb;
    } else if (i <= g.nx - 1 && j <= g.ny - 1) {
        const int o = g.at(i, j, k), sy = g.sy, sz = g.sz;
        const double u0 = x[o];
        const double ue = i < g.nx - 1 ? x[o + 1] : 0.0;
        const double uw = i > 2 ? x[o - 1] : u0;
        const double un = j < g.ny - 1 ? x[o + sy] : u0;
        const double us = j > 2 ? x[o - sy] : u0;
        const double ut = k < g.nz - 1 ? x[o + sz] : u0;
        const double ub = k > 2 ? x[o - sz] : u0;
        const double s = a[o] * (ue - u0) + a[o - 1] * (uw - u0) + b[o] * (un - u0) + b[o - sy] * (us - u0) +
                         c[o] * (ut - u0) + c[o - sz] * (ub - u0);
        const double bb = bvec[o];
        const double rr = bb - h2i * s;
        r[o] = rr;
        c1 = rr * rr;
        c2 = bb * bb;
    }
    double v1 = wave_sum(c1), v2 = wave_sum(c2);
    if ((tid & 63) == 0) { lds4[tid >> 6] = v1; lds4b[tid >> 6] = v2; }
    __syncthreads();
    if (tid == 0) {
        partial[2 * bidx] = lds4[0] + lds4[1] + lds4[2] + lds4[3];
        partial[2 * bidx + 1] = lds4b[0] + lds4b[1] + lds4b[2] + lds4b[3];
    }
}

__global__ void k_copy_cast(size_t n, const double *src, double *dst)
{
    size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p < n) dst[p] = src[p];
}

static int ensure_tol_buffers(Ctx &c)
{
    for (int l = 0; l < c.ng; l++) {
        Level &L = c.lev[l];
        const size_t n = L.g.n();
        if (c.mg_mode == 2 && !L.u32) {
            DG_HIP(hipMalloc(&L.a32, n * sizeof(float)));
            DG_HIP(hipMalloc(&L.b32, n * sizeof(float)));
            DG_HIP(hipMalloc(&L.c32, n * sizeof(float)));
            DG_HIP(hipMalloc(&L.u32, n * sizeof(float)));
            DG_HIP(hipMalloc(&L.f32, n * sizeof(float)));
            DG_HIP(hipMalloc(&L.r32, n * sizeof(float)));
            DG_HIP(hipMemset(L.u32, 0, n * sizeof(float)));
            DG_HIP(hipMemset(L.f32, 0, n * sizeof(float)));
            DG_HIP(hipMemset(L.r32, 0, n * sizeof(float)));
        }
        if (c.mg_mode == 1 && !L.u64) {
            DG_HIP(hipMalloc(&L.u64, n * sizeof(double)));
            DG_HIP(hipMalloc(&L.f64, n * sizeof(double)));
            DG_HIP(hipMalloc(&L.r64, n * sizeof(double)));
            DG_HIP(hipMemset(L.u64, 0, n * sizeof(double)));
            DG_HIP(hipMemset(L.f64, 0, n * sizeof(double)));
            DG_HIP(hipMemset(L.r64, 0, n * sizeof(double)));
        }
    }
    const size_t nfine = c.g.n();
    if (!c.pcg_r) {
        DG_HIP(hipMalloc(&c.pcg_r, nfine * sizeof(double)));
        DG_HIP(hipMalloc(&c.pcg_p, PCG_HIST * nfine * sizeof(double)));  // the direction ring
        DG_HIP(hipMalloc(&c.pcg_q, nfine * sizeof(double)));
        DG_HIP(hipMemset(c.pcg_r, 0, nfine * sizeof(double)));
        DG_HIP(hipMemset(c.pcg_p, 0, PCG_HIST * nfine * sizeof(double)));
        DG_HIP(hipMemset(c.pcg_q, 0, nfine * sizeof(double)));
    }
    const Grid g0 = c.lev[0].g;
    const size_t n0 = (size_t)(g0.nx - 2) * (g0.ny - 2) * (g0.nz - 2);
    if (n0 > 512) {
        set_error("coarsest multigrid level has %zu unknowns (limit 512)", n0);
        return -1;
    }
    // allocated once per context: the coarsest level's size is fixed by the
    // grid the context was created for
    if (!c.coarse_M) {
        DG_HIP(hipMalloc(&c.coarse_M, n0 * n0 * sizeof(double)));
        c.coarse_msize = n0 * n0;
    } else if (c.coarse_msize < n0 * n0) {
        set_error("coarse factor of %zu doubles is too small for %zu unknowns", c.coarse_msize, n0);
        return -1;
    }
    return 0;
}

// Enqueues one PCG iteration on c.stream (no host synchronisation): z = M r
// (one V-cycle), r.z and z.q_old, p_new = z + beta p_old into the
// direction ring and q = A p_new, p.q, r -= alpha q, r.r, scalar
// bookkeeping (alpha into the ring's step lengths), and the copy of r.r to
// pinned host memory.  x is updated by k_pcg_flush.
template <typename T>
static int pcg_iteration(Ctx &c)
{
    hipStream_t s = c.stream;
    Level &F = c.lev[c.ng - 1];
    const Grid g = F.g;
    const double h2i = lev_h2i(c, g);
    double *r = c.pcg_r, *q = c.pcg_q, *ds = c.dscal, *partial = c.red_partial;
    const size_t nslot = g.n();
    const dim3 ig = interior_grid(g), ib = interior_block();
    const int nblk = (int)(ig.x * ig.y * ig.z);
    // k_pcg_dir: 64 x PCG_TY columns of PCG_KCH planes
    const dim3 dgr((unsigned)((g.nx - 2 + 63) / 64), (unsigned)((g.ny - 2 + PCG_TY - 1) / PCG_TY),
                   (unsigned)((g.nz - 2 + PCG_KCH - 1) / PCG_KCH)),
        dib(64, PCG_TY);
    const int nblk_dir = (int)(dgr.x * dgr.y * dgr.z);
    T *u, *f, *rr_;
    const T *a_, *b_, *c_;
    level_ptrs<T>(F, u, f, rr_, a_, b_, c_);
    const double ncell = (double)(g.nx - 2) * (g.ny - 2) * (g.nz - 2);
    const bool split = rb_enabled(c);
    // search direction storage: FP32 with the FP32 V-cycle (mode 2), FP64
    // in the plain FP64 solver (mode 1)
    using PT = T;
    if (split) {
        // z = M r through the red-black split V-cycle; its right-hand side
        // (the split copy of r) was written by the previous update or by the
        // solve's set-up
        if (rb_vcycle(c, nullptr)) return -1;
        const ZSplit<T> z = rb_top_z<T>(c);
        hipLaunchKernelGGL(k_dot_rz_zq<ZSplit<T>>, dim3(RED_BLOCKS), dim3(256), 0, s, g, r, z, q, partial);
        DG_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_final2, dim3(1), dim3(256), 0, s, partial, RED_BLOCKS, ds + S_RZ);
        DG_CHECK_LAUNCH();
        prof_begin(PK_PCG_DIR, s);
        hipLaunchKernelGGL((k_pcg_dir<ZSplit<T>, PT>), dgr, dib, 0, s, g, z, (PT *)c.pcg_p, nslot, q, F.a, F.b, F.c, h2i,
                           ds, partial);
        prof_end(PK_PCG_DIR, s, ((double)sizeof(T) + 2.0 * sizeof(PT) + 4.0 * 8.0) * ncell);
        DG_CHECK_LAUNCH();
    } else {
        if (sizeof(T) == 4)
            hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((g.n() + 255) / 256)), dim3(256), 0, s, g.n(), r, (float *)f);
        else
            hipLaunchKernelGGL(k_copy_cast, dim3((unsigned)((g.n() + 255) / 256)), dim3(256), 0, s, g.n(), r,
                               (double *)f);
        DG_CHECK_LAUNCH();
        if (vcycle<T>(c, c.ng - 1)) return -1;
        const ZNatural<T> z{(const T *)u, g};
        hipLaunchKernelGGL(k_dot_rz_zq<ZNatural<T>>, dim3(RED_BLOCKS), dim3(256), 0, s, g, r, z, q, partial);
        DG_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_final2, dim3(1), dim3(256), 0, s, partial, RED_BLOCKS, ds + S_RZ);
        DG_CHECK_LAUNCH();
        prof_begin(PK_PCG_DIR, s);
        hipLaunchKernelGGL((k_pcg_dir<ZNatural<T>, PT>), dgr, dib, 0, s, g, z, (PT *)c.pcg_p, nslot, q, F.a, F.b, F.c,
                           h2i, ds, partial);
        prof_end(PK_PCG_DIR, s, ((double)sizeof(T) + 2.0 * sizeof(PT) + 4.0 * 8.0) * ncell);
        DG_CHECK_LAUNCH();
    }
    // p.q: k_pcg_dir's per-workgroup partials summed in one launch
    hipLaunchKernelGGL(k_sum_wide, dim3(1), dim3(1024), 0, s, partial, nblk_dir, ds + S_PQ);
    DG_CHECK_LAUNCH();
    prof_begin(PK_PCG_UPDATE, s);
    hipLaunchKernelGGL(k_pcg_update<PT>, ig, ib, 0, s, g, r, q, ds, partial, split ? rb_f_of<T>(F) : nullptr,
                       rb_geom(g));
    prof_end(PK_PCG_UPDATE, s, (3.0 * 8.0 + (split ? (double)sizeof(T) : 0.0)) * ncell);
    DG_CHECK_LAUNCH();
    // r.r: chunk sums, then their sum with the scalar bookkeeping
    {
        const int nchunk = (nblk + RP_CHUNK - 1) / RP_CHUNK;
        double *chunks = partial + 2 * (size_t)nblk;
        hipLaunchKernelGGL(k_partial_chunks, dim3((unsigned)nchunk), dim3(256), 0, s, partial, nblk, 1, chunks);
        DG_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_final_scalars, dim3(1), dim3(256), 0, s, chunks, nchunk, ds, c.dpoll);
        DG_CHECK_LAUNCH();
    }
    return 0;
}

// The iteration's kernel sequence and arguments depend only on the level
// layout, the buffers and the solver options, so it is captured once per
// (mode, nu1, nu2, fused) and replayed: one graph launch per iteration
// instead of ~10 launches per level.  The kernels take the direction
// ring's slots from the device iteration count, so one graph serves every
// iteration.
template <typename T>
static int pcg_graphs(Ctx &c)
{
    const int rb = rb_enabled(c);
    // everything that changes the captured kernel sequence or its arguments
    const int key[13] = {c.mg_mode, c.mg_nu1, c.mg_nu2, c.ng, (int)c.g.nx,
                         rb, rb ? c.lev[c.ng - 1].rb_ct : -1, rb ? c.rb_tail : -1, c.mg_nuc,
                         (int)llround(c.mg_omega * 1e9), (int)c.g.nz, rb ? c.rb_tiles : -1,
                         rb ? c.rb_tile_nu : -1};
    bool ok = c.pcg_graph[0] != nullptr;
    for (int q = 0; q < 13 && ok; q++) ok = c.pcg_graph_key[q] == key[q];
    if (ok) return 0;
    for (int q = 0; q < 2; q++) {
        if (c.pcg_graph[q]) {
            (void)hipGraphExecDestroy(c.pcg_graph[q]);
            c.pcg_graph[q] = nullptr;
        }
    }
    {
        hipGraph_t graph;
        DG_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
        int rc = pcg_iteration<T>(c);
        hipError_t e = hipStreamEndCapture(c.stream, &graph);
        if (rc) return -1;
        if (e != hipSuccess) {
            set_error("PCG graph capture: %s", hipGetErrorString(e));
            return -1;
        }
        DG_HIP(hipGraphInstantiate(&c.pcg_graph[0], graph, nullptr, nullptr, 0));
        DG_HIP(hipGraphDestroy(graph));
    }
    for (int q = 0; q < 13; q++) c.pcg_graph_key[q] = key[q];
    return 0;
}

template <typename T>
static int pcg_solve(Ctx &c)
{
    hipStream_t s = c.stream;
    Level &F = c.lev[c.ng - 1];
    const Grid g = F.g;
    const double h2i = lev_h2i(c, g);
    double *x = F.u, *bvec = F.rhs, *r = c.pcg_r;
    double *ds = c.dscal;
    const dim3 ig = interior_grid(g), ib = interior_block();
    const int nblk = (int)(ig.x * ig.y * ig.z);
    double *partial = c.red_partial;
    // guess 1: the previous pressure (residual-checked below); guess 0 and
    // the learned guess 2: zero (the network predicts the pressure itself)
    if (c.mg_guess == 1) {
        DG_HIP(hipMemcpyAsync(x, c.pprev, g.n() * sizeof(double), hipMemcpyDeviceToDevice, s));
    } else {
        if (mg_fill(g.n(), s, x, 0.0)) return -1;
    }
    // the split V-cycle's first right-hand side: written with r = b from a
    // zero guess, loaded below otherwise
    const bool split = rb_enabled(c);
    if (c.mg_guess == 1)
        hipLaunchKernelGGL((k_init_resid<false, T>), ig, ib, 0, s, g, r, x, bvec, F.a, F.b, F.c, h2i, partial,
                           (T *)nullptr, rb_geom(g));
    else
        hipLaunchKernelGGL((k_init_resid<true, T>), ig, ib, 0, s, g, r, x, bvec, F.a, F.b, F.c, h2i, partial,
                           split ? rb_f_of<T>(F) : nullptr, rb_geom(g));
    DG_CHECK_LAUNCH();
    if (sum_partials(s, partial, nblk, 2, partial + 2 * (size_t)nblk, ds + S_RR)) return -1;  // r.r and b.b
    if (c.mg_guess == 1) {
        // residual check of the previous-step guess against the zero guess
        hipLaunchKernelGGL(k_guess_check, ig, ib, 0, s, g, x, r, bvec, ds);
        DG_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_guess_check_scalar, dim3(1), dim3(1), 0, s, ds);
        DG_CHECK_LAUNCH();
    }
    // learned guess (mg_guess 2) and sample capture
    const bool hooks = guess_wanted(c);
    bool r_changed = false;
    if (hooks && guess_solve_begin(c, x, r, h2i, c.mg_guess == 2, ds + S_RR, &r_changed)) return -1;
    if (split && (c.mg_guess == 1 || r_changed) && rb_load_rhs(c, r)) return -1;
    // coarse factorisation for this solve, FIRST flag
    {
        Level &L0 = c.lev[0];
        hipLaunchKernelGGL(k_coarse_factor, dim3(1), dim3(256), 0, s, L0.g, L0.a, L0.b, L0.c, lev_h2i(c, L0.g),
                           c.coarse_M, ds + S_CINFO);
        DG_CHECK_LAUNCH();
        c.hscal[S_FIRST] = 1.0;
        c.hscal[S_IT] = 0.0;
        c.hscal[S_STOP] = 0.0;
        c.hscal[S_TOL] = c.mg_tol;
        DG_HIP(hipMemcpyAsync(ds + S_FIRST, c.hscal + S_FIRST, 4 * sizeof(double), hipMemcpyHostToDevice, s));
    }
    // r.r, b.b (S_RR, S_BB) and the factorisation's info (S_CINFO)
    static_assert(S_CINFO == S_RR + 6, "one copy covers r.r .. the coarse factor's info");
    DG_HIP(hipMemcpyAsync(c.hscal + S_RR, ds + S_RR, 7 * sizeof(double), hipMemcpyDeviceToHost, s));
    DG_HIP(hipStreamSynchronize(s));
    double bnorm = sqrt(c.hscal[S_RR + 1]);
    if (bnorm == 0.0) bnorm = 1.0;
    double rel = sqrt(c.hscal[S_RR]) / bnorm;
    c.mg_iters_last = 0;
    c.mg_relres_last = rel;
    if (!std::isfinite(c.hscal[S_RR + 1]) || !std::isfinite(rel)) {
        set_error("pressure solve: non-finite right-hand side or initial residual (|b|^2 = %g, |r|/|b| = %g)",
                  c.hscal[S_RR + 1], rel);
        return -1;
    }
    if (c.hscal[S_CINFO] != 0.0) {
        set_error("pressure solve: coarsest-level Cholesky factorisation failed at pivot %d",
                  (int)c.hscal[S_CINFO]);
        return -1;
    }
    const bool graphs = c.use_graphs && prof().kernel == PK_NONE;
    if (graphs && pcg_graphs<T>(c)) return -1;
    // x += alpha_j p_j over the ring's first count slots
    using PT = T;
    auto flush = [&](int count) {
        return launch_flush<PT>(count, s, g, x, (const PT *)c.pcg_p, g.n(), ds);
    };
    // the iteration's r.r arrives in mapped host memory: poll for its count
    // (bounded), falling back to a stream synchronisation
    volatile double *poll = c.hpoll;
    poll[1] = 0.0;
    auto wait_iteration = [&](int n) {
        const auto t0 = std::chrono::steady_clock::now();
        while (poll[1] < (double)n)
            if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(2)) return false;
        return true;
    };
    // iterations queued (launched) and finished with their residual read (it)
    int it = 0, launched = 0;
    auto launch = [&]() {
        if (graphs) {
            if (hipGraphLaunch(c.pcg_graph[0], s) != hipSuccess) {
                set_error("PCG graph launch failed");
                return -1;
            }
        } else {
            if (pcg_iteration<T>(c)) return -1;
        }
        launched++;
        if (launched % PCG_HIST == 0 && flush(PCG_HIST)) {
            set_error("k_pcg_flush launch failed");
            return -1;
        }
        return 0;
    };
    bool stop = !(rel >= c.mg_tol);
    while (!stop && it < c.mg_maxit) {
        if (launched == it && launch()) return -1;
        // look ahead while the residual is far above the tolerance; never
        // to an iteration that fills the ring, whose flush would take that
        // iteration's slot whether it ran or was skipped
        if (launched == it + 1 && launched < c.mg_maxit && (launched + 1) % PCG_HIST != 0 &&
            c.pcg_lookahead > 0.0 && rel > c.pcg_lookahead * c.mg_tol && launch())
            return -1;
        if (!wait_iteration(it + 1)) {
            DG_HIP(hipStreamSynchronize(s));
            if (poll[1] < (double)(it + 1)) {
                set_error("PCG iteration %d reported no residual", it + 1);
                return -1;
            }
        }
        it++;
        const volatile double *pair = poll + 2 + 2 * (it & 1);
        rel = sqrt(pair[0]) / bnorm;
        stop = pair[1] != 0.0;
    }
    if (it % PCG_HIST != 0 && flush(it % PCG_HIST)) {
        set_error("k_pcg_flush launch failed");
        return -1;
    }
    c.mg_iters_last = it;
    c.mg_relres_last = rel;
    if (!std::isfinite(rel)) {
        set_error("pressure solve: non-finite residual after %d PCG iterations", it);
        return -1;
    }
    if (!(rel < c.mg_tol)) {
        set_error("pressure solve: |r|/|b| = %.3e after %d PCG iterations (maxit %d) has not reached tol %.3e", rel,
                  it, c.mg_maxit, c.mg_tol);
        return -1;
    }
    if (op_bc_scalar(g, s, x)) return -1;
    if (hooks && guess_solve_end(c, x, h2i, true)) return -1;
    DG_HIP(hipMemcpyAsync(c.pprev, x, g.n() * sizeof(double), hipMemcpyDeviceToDevice, s));
    return 0;
}

int op_mg_solve_tol(Ctx &c)
{
    hipStream_t s = c.stream;
    if (ensure_tol_buffers(c)) return -1;
    const Grid g = c.g;
    (void)g;  // the partial-sum buffer is sized for this grid by alloc_levels
    // with split storage the coarse levels' split coefficients are written
    // by acfromac itself (the arrays are allocated first)
    const bool split = rb_enabled(c);
    if (split && rb_alloc(c)) return -1;
    for (int gi = c.ng - 1; gi >= 1; gi--) {
        Level &C = c.lev[gi - 1], &F = c.lev[gi];
        const bool cs = split && gi - 1 >= 1;
        if (mg_acfromac(C.g, s, C.a, C.b, C.c, C.ei, F.a, F.b, F.c, cs ? C.rb_coef : nullptr, cs ? C.rb_ct : -1))
            return -1;
    }
    if (split) {
        c.coarse_coef_ready = 1;
        if (rb_prepare(c)) return -1;
    } else if (c.mg_mode == 2) {
        for (int l = 0; l < c.ng; l++) {
            Level &L = c.lev[l];
            const size_t n = L.g.n();
            const unsigned nb = (unsigned)((n + 255) / 256);
            hipLaunchKernelGGL(k_to_f32, dim3(nb), dim3(256), 0, s, n, L.a, L.a32);
            hipLaunchKernelGGL(k_to_f32, dim3(nb), dim3(256), 0, s, n, L.b, L.b32);
            hipLaunchKernelGGL(k_to_f32, dim3(nb), dim3(256), 0, s, n, L.c, L.c32);
            DG_CHECK_LAUNCH();
        }
    }
    int (*solve)(Ctx &) = c.mg_mode == 2 ? pcg_solve<float> : pcg_solve<double>;
    if (guess_eval_enabled()) return guess_eval_solves(c, solve);
    return solve(c);
}

}  // namespace dg
