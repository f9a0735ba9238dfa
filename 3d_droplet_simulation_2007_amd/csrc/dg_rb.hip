// dg_rb.hip -- the FP32 V-cycle of tolerance mode 2 on red-black split
// storage (layout in dg_rb.h).  Same algorithm as the natural-layout
// V-cycle in dg_mg.hip (symmetric red-black Gauss-Seidel V(nu1,nu2) with
// the operator's own diagonal, 8-child average restriction, injection
// prolongation, acfromac coarse operators, FP64 Cholesky on the coarsest
// level); the coefficients may be rounded to FP16.

#include <hip/hip_runtime.h>

#include "dg_coarse.h"
#include "dg_prof.h"
#include "dg_rb.h"
#include "dg_util.h"

namespace dg {

int coarse_solve_f32(Ctx &c);     // dg_mg.hip
const double *coarse_factor();  // dg_mg.hip: Cholesky factor of the coarsest level

namespace {

// three coefficient planes (a, b, c), each 2*half entries in split order
template <typename CT>
struct Coef {
    const CT *p;
    size_t n2;  // 2 * half
    __device__ inline float a(size_t e) const { return (float)p[e]; }
    __device__ inline float b(size_t e) const { return (float)p[n2 + e]; }
    __device__ inline float c(size_t e) const { return (float)p[2 * n2 + e]; }
};

// thread (i, j, k) of the interior, k = blockIdx.z
template <typename CT>
__global__ void __launch_bounds__(256) k_rb_coef(Grid g, RbGeom r, const double *a, const double *b, const double *c,
                                                 CT *out)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 2;
    const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) + 2;
    const int k = (int)blockIdx.z + 2;
    if (i > g.nx - 1 || j > g.ny - 1) return;
    const size_t n2 = 2 * r.half;
    const int o = g.at(i, j, k);
    const size_t e = r.at(i, j, k);
    out[e] = (CT)a[o];
    out[n2 + e] = (CT)b[o];
    out[2 * n2 + e] = (CT)c[o];
}

// finest-level right-hand side: f (split FP32) = r (natural FP64)
__global__ void __launch_bounds__(256) k_rb_in(Grid g, RbGeom r, const double *src, float *dst)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 2;
    const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) + 2;
    const int k = (int)blockIdx.z + 2;
    if (i > g.nx - 1 || j > g.ny - 1) return;
    dst[r.at(i, j, k)] = (float)src[g.at(i, j, k)];
}

// One Gauss-Seidel colour pass: thread (m, j, k) updates the colour's cell
// i = 2 + ((colour+j+k)&1) + 2m of row (j,k).  Neighbours (all of the other
// colour) and the other colour's face coefficients (west, south, bottom) sit
// in the other half at m+p-1 / m+p (x) or m (y, z).  Boundary folding as the
// natural-layout pass: a missing x-min / y / z neighbour drops its face, the
// x-max neighbour is 0 with its face kept in the diagonal.  ZERO: the
// iterate is zero on entry (first pass of a V-cycle level), u is not read.
// TOP only separates the finest level's launches into their own symbol.
template <typename CT, bool ZERO>
__device__ inline void rb_cell(const RbGeom &r, float *u, const float *f, const Coef<CT> &cf, float h2, int colour,
                               float om, int zown, int m, int j, int k)
{
    const int p = (colour + j + k) & 1;
    const int i = 2 + p + 2 * m;
    const size_t row = ((size_t)(k - 2) * r.my + (j - 2)) * r.mh;
    const size_t own = (size_t)colour * r.half + row + m;
    const size_t oth = (size_t)(1 - colour) * r.half;
    const size_t rs = (size_t)r.mh, ps = (size_t)r.mh * r.my;  // row and plane strides of one colour
    const float ae = cf.a(own), bn = cf.b(own), ct = cf.c(own);
    const bool hw = i > 2, hn = j < r.ny - 1, hs = j > 2, ht = k < r.nz - 1, hb = k > 2;
    const float aw = hw ? cf.a(oth + row + m + p - 1) : 0.f;
    const float bs = hs ? cf.b(oth + row - rs + m) : 0.f;
    const float cb = hb ? cf.c(oth + row - ps + m) : 0.f;
    float s = 0.f, d = ae;
    if (!ZERO) {
        if (i < r.nx - 1) s += ae * u[oth + row + m + p];
        if (hw) s += aw * u[oth + row + m + p - 1];
        if (hn) s += bn * u[oth + row + rs + m];
        if (hs) s += bs * u[oth + row - rs + m];
        if (ht) s += ct * u[oth + row + ps + m];
        if (hb) s += cb * u[oth + row - ps + m];
    }
    if (hw) d += aw;
    if (hn) d += bn;
    if (hs) d += bs;
    if (ht) d += ct;
    if (hb) d += cb;
    const float gsv = (s - h2 * f[own]) / d;
    const float uold = (ZERO || zown || om == 1.f) ? 0.f : u[own];  // Gauss-Seidel needs no old value
    u[own] = om * gsv + (1.f - om) * uold;
}

template <typename CT, bool ZERO, bool TOP>
__global__ void __launch_bounds__(256) k_rb_colour(RbGeom r, float *__restrict__ u, const float *__restrict__ f,
                                                   Coef<CT> cf, float h2, int colour, float om, int zown)
{
    const int m = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) + 2;
    const int k = (int)blockIdx.z + 2;
    if (m >= r.mh || j > r.ny - 1) return;
    rb_cell<CT, ZERO>(r, u, f, cf, h2, colour, om, zown, m, j, k);
}

// four consecutive entries (16-byte aligned for FP32, 8-byte for FP16)
__device__ inline void ld4(const float *p, float v[4])
{
    const float4 q = *reinterpret_cast<const float4 *>(p);
    v[0] = q.x;
    v[1] = q.y;
    v[2] = q.z;
    v[3] = q.w;
}
typedef _Float16 h4 __attribute__((ext_vector_type(4)));
__device__ inline void ld4(const _Float16 *p, float v[4])
{
    const h4 q = *reinterpret_cast<const h4 *>(p);
    v[0] = (float)q[0];
    v[1] = (float)q[1];
    v[2] = (float)q[2];
    v[3] = (float)q[3];
}

// The colour pass with four consecutive cells of a row per thread (rows of
// a multiple of 4 cells): one vector load per neighbour row instead of
// four scalar loads, same arithmetic in the same order per cell.
template <typename CT, bool ZERO, bool TOP>
__global__ void __launch_bounds__(256) k_rb_colour4(RbGeom r, float *__restrict__ u, const float *__restrict__ f,
                                                    Coef<CT> cf, float h2, int colour, float om, int zown)
{
    const int m0 = 4 * (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) + 2;
    const int k = (int)blockIdx.z + 2;
    if (m0 >= r.mh || j > r.ny - 1) return;
    const int p = (colour + j + k) & 1;
    const size_t row = ((size_t)(k - 2) * r.my + (j - 2)) * r.mh;
    const size_t own = (size_t)colour * r.half + row + m0;
    const size_t oth = (size_t)(1 - colour) * r.half + row + m0;  // other colour, same row, m0
    const size_t rs = (size_t)r.mh, ps = (size_t)r.mh * r.my;
    const bool hn = j < r.ny - 1, hs = j > 2, ht = k < r.nz - 1, hb = k > 2;
    float ae[4], bn[4], ct[4], aw[4] = {0.f, 0.f, 0.f, 0.f}, bs[4] = {0.f, 0.f, 0.f, 0.f},
                                 cb[4] = {0.f, 0.f, 0.f, 0.f};
    ld4(cf.p + own, ae);
    ld4(cf.p + cf.n2 + own, bn);
    ld4(cf.p + 2 * cf.n2 + own, ct);
    // west faces: the other colour's a at m+p-1
    {
        float t[4];
        ld4(cf.p + oth, t);
        if (p) {
            for (int q = 0; q < 4; q++) aw[q] = t[q];
        } else {
            aw[0] = m0 > 0 ? cf.a(oth - 1) : 0.f;
            aw[1] = t[0];
            aw[2] = t[1];
            aw[3] = t[2];
        }
    }
    if (hs) ld4(cf.p + cf.n2 + oth - rs, bs);
    if (hb) ld4(cf.p + 2 * cf.n2 + oth - ps, cb);
    float ue[4] = {0.f, 0.f, 0.f, 0.f}, uw[4] = {0.f, 0.f, 0.f, 0.f}, un[4], us[4], ut[4], ub[4];
    if (!ZERO) {
        float t[4];
        ld4(u + oth, t);
        if (p) {  // west m .. , east m+1 ..
            for (int q = 0; q < 4; q++) uw[q] = t[q];
            ue[0] = t[1];
            ue[1] = t[2];
            ue[2] = t[3];
            ue[3] = m0 + 4 < r.mh ? u[oth + 4] : 0.f;
        } else {  // west m-1 .., east m ..
            for (int q = 0; q < 4; q++) ue[q] = t[q];
            uw[0] = m0 > 0 ? u[oth - 1] : 0.f;
            uw[1] = t[0];
            uw[2] = t[1];
            uw[3] = t[2];
        }
        if (hn) ld4(u + oth + rs, un);
        if (hs) ld4(u + oth - rs, us);
        if (ht) ld4(u + oth + ps, ut);
        if (hb) ld4(u + oth - ps, ub);
    }
    float fv[4], out[4], uo[4] = {0.f, 0.f, 0.f, 0.f};
    ld4(f + own, fv);
    if (!ZERO && !zown && om != 1.f) ld4(u + own, uo);  // Gauss-Seidel needs no old value
    for (int q = 0; q < 4; q++) {
        const int i = 2 + p + 2 * (m0 + q);
        const bool hw = i > 2, he = i < r.nx - 1;
        float s = 0.f, d = ae[q];
        if (!ZERO) {
            if (he) s += ae[q] * ue[q];
            if (hw) s += aw[q] * uw[q];
            if (hn) s += bn[q] * un[q];
            if (hs) s += bs[q] * us[q];
            if (ht) s += ct[q] * ut[q];
            if (hb) s += cb[q] * ub[q];
        }
        if (hw) d += aw[q];
        if (hn) d += bn[q];
        if (hs) d += bs[q];
        if (ht) d += ct[q];
        if (hb) d += cb[q];
        const float gsv = (s - h2 * fv[q]) / d;
        out[q] = om * gsv + (1.f - om) * uo[q];
    }
    *reinterpret_cast<float4 *>(u + own) = make_float4(out[0], out[1], out[2], out[3]);
}

// residual of one fine cell (f - A u, boundary folding of the natural resid)
template <typename CT>
__device__ inline float rb_resid_cell(const RbGeom &r, const float *u, const float *f, const Coef<CT> &cf, float h2i,
                                      int i, int j, int k)
{
    auto at = [&](int ii, int jj, int kk) { return r.at32(ii, jj, kk); };
    const unsigned e = at(i, j, k);
    const float u0 = u[e];
    const float ue = i < r.nx - 1 ? u[at(i + 1, j, k)] : 0.f;
    const float uw = i > 2 ? u[at(i - 1, j, k)] : u0;
    const float un = j < r.ny - 1 ? u[at(i, j + 1, k)] : u0;
    const float us = j > 2 ? u[at(i, j - 1, k)] : u0;
    const float ut = k < r.nz - 1 ? u[at(i, j, k + 1)] : u0;
    const float ub = k > 2 ? u[at(i, j, k - 1)] : u0;
    const float aw = i > 2 ? cf.a(at(i - 1, j, k)) : 0.f;
    const float bs = j > 2 ? cf.b(at(i, j - 1, k)) : 0.f;
    const float cb = k > 2 ? cf.c(at(i, j, k - 1)) : 0.f;
    const float s = cf.a(e) * (ue - u0) + aw * (uw - u0) + cf.b(e) * (un - u0) + bs * (us - u0) + cf.c(e) * (ut - u0) +
                    cb * (ub - u0);
    return f[e] - h2i * s;
}

// coarse right-hand side = average of the 8 children's residuals; the
// coarse level is split (fcs) or the natural-layout coarsest level (fcn)
template <typename CT>
__device__ inline void rb_restrict_cell(const RbGeom &rf, const float *u, const float *f, const Coef<CT> &cf,
                                        float h2i, const RbGeom &rc, float *fcs, const Grid &gc, float *fcn, int I,
                                        int J, int K)
{
    float t = 0.f;
    for (int dz = 0; dz < 2; dz++)
        for (int dy = 0; dy < 2; dy++)
            for (int dx = 0; dx < 2; dx++)
                t += rb_resid_cell<CT>(rf, u, f, cf, h2i, 2 * I + 2 + dx, 2 * J + 2 + dy, 2 * K + 2 + dz);
    const float v = 0.125f * t;
    if (fcs)
        fcs[rc.at(I + 2, J + 2, K + 2)] = v;
    else
        fcn[gc.at(I + 2, J + 2, K + 2)] = v;
}

// thread (I, J, K) of the coarse interior, K = blockIdx.z
template <typename CT>
__global__ void __launch_bounds__(256) k_rb_resid_restrict(RbGeom rf, const float *u, const float *f, Coef<CT> cf,
                                                           float h2i, RbGeom rc, float *fcs, Grid gc, float *fcn)
{
    const int I = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int J = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    const int K = (int)blockIdx.z;
    if (I >= gc.nx - 2 || J >= gc.ny - 2) return;
    rb_restrict_cell<CT>(rf, u, f, cf, h2i, rc, fcs, gc, fcn, I, J, K);
}

// fine u += coarse u of the parent cell (injection), fine in split order:
// entry m of row (j, k) in both colours, cells i = 2 + p + 2m with p = 0, 1,
// both children of coarse (m+2, j', k')
__device__ inline void rb_prolong_entry(const RbGeom &rf, float *u, const RbGeom &rc, const float *ucs,
                                        const Grid &gc, const float *ucn, int m, int j, int k)
{
    const int ic = m + 2, jc = (j - 2) / 2 + 2, kc = (k - 2) / 2 + 2;
    const float e = ucs ? ucs[rc.at(ic, jc, kc)] : ucn[gc.at(ic, jc, kc)];
    const size_t o = ((size_t)(k - 2) * rf.my + (j - 2)) * rf.mh + m;
    u[o] += e;
    u[rf.half + o] += e;
}

__global__ void __launch_bounds__(256) k_rb_prolong(RbGeom rf, float *u, RbGeom rc, const float *ucs, Grid gc,
                                                    const float *ucn)
{
    const int m = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) + 2;
    const int k = (int)blockIdx.z + 2;
    if (m >= rf.mh || j > rf.ny - 1) return;
    rb_prolong_entry(rf, u, rc, ucs, gc, ucn, m, j, k);
}

// The V-cycle below a small level in one workgroup: every colour pass,
// restriction, the coarsest solve and every prolongation of the levels
// 1..top as loops of TAIL_THREADS threads separated by barriers, with the
// per-cell arithmetic of the kernels above (same results, ~4 launches per
// level fewer).  Levels are indexed as in the context (0 = coarsest, the
// natural-layout direct solve).
constexpr int TAIL_THREADS = 1024;
constexpr int TAIL_MAX_LEVELS = 8;
constexpr size_t TAIL_MAX_CELLS = 4096;  // largest level taken into the tail

template <typename CT>
struct RbTail {
    int top;  // highest level in the tail (>= 1)
    RbGeom r[TAIL_MAX_LEVELS];
    float *u[TAIL_MAX_LEVELS], *f[TAIL_MAX_LEVELS];
    Coef<CT> cf[TAIL_MAX_LEVELS];
    float h2[TAIL_MAX_LEVELS], h2i[TAIL_MAX_LEVELS];
    int nu1[TAIL_MAX_LEVELS], nu2[TAIL_MAX_LEVELS];
    float om;
    Grid g0;  // coarsest level (natural layout)
    float *f0, *u0;
    const double *M;
};

template <typename CT, bool ZERO>
__device__ inline void tail_pass(const RbTail<CT> &t, int l, int colour, int zown)
{
    const RbGeom &r = t.r[l];
    const int n = (int)r.half;
    for (int e = (int)threadIdx.x; e < n; e += TAIL_THREADS) {
        const int m = e % r.mh, q = e / r.mh;
        rb_cell<CT, ZERO>(r, t.u[l], t.f[l], t.cf[l], t.h2[l], colour, t.om, zown, m, q % r.my + 2, q / r.my + 2);
    }
    __syncthreads();
}

template <typename CT>
__global__ void __launch_bounds__(TAIL_THREADS) k_rb_tail(RbTail<CT> t)
{
    __shared__ double y[512];
    for (int l = t.top; l >= 1; l--) {
        for (int q = 0; q < t.nu1[l]; q++) {
            if (q == 0)
                tail_pass<CT, true>(t, l, 0, 1);
            else
                tail_pass<CT, false>(t, l, 0, 0);
            tail_pass<CT, false>(t, l, 1, q == 0);
        }
        const RbGeom &rf = t.r[l];
        const bool csplit = l - 1 >= 1;
        const int cx = rf.mh, cy = rf.my / 2, n = cx * cy * (rf.mz / 2);  // coarse interior
        for (int e = (int)threadIdx.x; e < n; e += TAIL_THREADS) {
            const int I = e % cx, q = e / cx;
            rb_restrict_cell<CT>(rf, t.u[l], t.f[l], t.cf[l], t.h2i[l], t.r[csplit ? l - 1 : l],
                                 csplit ? t.f[l - 1] : nullptr, t.g0, csplit ? nullptr : t.f0, I, q % cy, q / cy);
        }
        __syncthreads();
    }
    coarse_subst<float>(t.g0, t.M, t.f0, t.u0, y, (int)threadIdx.x, TAIL_THREADS);
    __syncthreads();
    for (int l = 1; l <= t.top; l++) {
        const RbGeom &rf = t.r[l];
        const bool csplit = l - 1 >= 1;
        const int n = (int)rf.half;
        for (int e = (int)threadIdx.x; e < n; e += TAIL_THREADS) {
            const int m = e % rf.mh, q = e / rf.mh;
            rb_prolong_entry(rf, t.u[l], csplit ? t.r[l - 1] : rf, csplit ? t.u[l - 1] : nullptr, t.g0,
                             csplit ? nullptr : t.u0, m, q % rf.my + 2, q / rf.my + 2);
        }
        __syncthreads();
        for (int q = 0; q < t.nu2[l]; q++) {
            tail_pass<CT, false>(t, l, 1, 0);
            tail_pass<CT, false>(t, l, 0, 0);
        }
    }
}

// 256-thread block shaped to a row of `lanes` threads: x extent the
// smallest power of two in 8..64 covering it, the rest along y
dim3 row_block(int lanes)
{
    int bx = 64;
    while (bx > 8 && bx / 2 >= lanes) bx /= 2;
    return dim3((unsigned)bx, (unsigned)(256 / bx));
}

dim3 row_grid(const dim3 &b, int lanes, int rows, int planes)
{
    return dim3((unsigned)((lanes + (int)b.x - 1) / (int)b.x), (unsigned)((rows + (int)b.y - 1) / (int)b.y),
                (unsigned)planes);
}

double lev_h2i(const Ctx &c, const Grid &g)
{
    const double h = c.s.lengthsize / (double)(g.nx - 2);
    return 1.0 / (h * h);
}

// zero: the iterate is zero on entry; zown: this pass's own cells are zero
// (the second pass of the first sweep)
template <typename CT>
int colour_pass(Ctx &c, int l, int colour, bool zero, int zown)
{
    const float om = (float)c.mg_omega;
    Level &L = c.lev[l];
    const RbGeom r = rb_geom(L.g);
    const Coef<CT> cf{(const CT *)L.rb_coef, 2 * r.half};
    const float h2 = (float)(1.0 / lev_h2i(c, L.g));
    const dim3 b1 = row_block(r.mh), grid = row_grid(b1, r.mh, r.my, r.mz);
    const bool top = l == c.ng - 1;
    // algorithmic bytes: own f read and u write (one colour, FP32), the other
    // colour's u (not read from a zero iterate) and all coefficients
    const double cells = (double)r.half;
    const double bytes = cells * (zero ? 8.0 : (om == 1.f || zown ? 12.0 : 16.0)) + 6.0 * cells * (double)sizeof(CT);
    if (top) prof_begin(PK_RB_COLOUR, c.stream);
    if (r.mh % 4 == 0) {
        const dim3 b4 = row_block(r.mh / 4), g4 = row_grid(b4, r.mh / 4, r.my, r.mz);
        if (top && zero)
            hipLaunchKernelGGL((k_rb_colour4<CT, true, true>), g4, b4, 0, c.stream, r, L.rb_u, L.rb_f, cf, h2, colour, om, zown);
        else if (top)
            hipLaunchKernelGGL((k_rb_colour4<CT, false, true>), g4, b4, 0, c.stream, r, L.rb_u, L.rb_f, cf, h2, colour, om, zown);
        else if (zero)
            hipLaunchKernelGGL((k_rb_colour4<CT, true, false>), g4, b4, 0, c.stream, r, L.rb_u, L.rb_f, cf, h2, colour, om, zown);
        else
            hipLaunchKernelGGL((k_rb_colour4<CT, false, false>), g4, b4, 0, c.stream, r, L.rb_u, L.rb_f, cf, h2, colour, om, zown);
    } else if (top && zero)
        hipLaunchKernelGGL((k_rb_colour<CT, true, true>), grid, b1, 0, c.stream, r, L.rb_u, L.rb_f, cf, h2, colour, om, zown);
    else if (top)
        hipLaunchKernelGGL((k_rb_colour<CT, false, true>), grid, b1, 0, c.stream, r, L.rb_u, L.rb_f, cf, h2, colour, om, zown);
    else if (zero)
        hipLaunchKernelGGL((k_rb_colour<CT, true, false>), grid, b1, 0, c.stream, r, L.rb_u, L.rb_f, cf, h2, colour, om, zown);
    else
        hipLaunchKernelGGL((k_rb_colour<CT, false, false>), grid, b1, 0, c.stream, r, L.rb_u, L.rb_f, cf, h2, colour, om, zown);
    if (top) prof_end(PK_RB_COLOUR, c.stream, bytes);
    DG_CHECK_LAUNCH();
    return 0;
}

// highest level taken into the single-workgroup tail, or -1
int tail_level(const Ctx &c)
{
    int lt = -1;
    for (int l = 1; l < c.ng - 1 && l < TAIL_MAX_LEVELS; l++)
        if (2 * rb_geom(c.lev[l].g).half <= TAIL_MAX_CELLS) lt = l;
    return lt;
}

template <typename CT>
int launch_tail(Ctx &c, int lt)
{
    const Grid &g0 = c.lev[0].g;
    if ((g0.nx - 2) * (g0.ny - 2) * (g0.nz - 2) > 512) {
        set_error("split V-cycle tail: coarsest level above 512 unknowns");
        return -1;
    }
    RbTail<CT> t{};
    t.top = lt;
    for (int l = 1; l <= lt; l++) {
        Level &L = c.lev[l];
        t.r[l] = rb_geom(L.g);
        t.u[l] = L.rb_u;
        t.f[l] = L.rb_f;
        t.cf[l] = Coef<CT>{(const CT *)L.rb_coef, 2 * t.r[l].half};
        t.h2[l] = (float)(1.0 / lev_h2i(c, L.g));
        t.h2i[l] = (float)lev_h2i(c, L.g);
        t.nu1[l] = c.mg_nuc <= 0 ? c.mg_nu1 : c.mg_nuc;
        t.nu2[l] = c.mg_nuc <= 0 ? c.mg_nu2 : c.mg_nuc;
    }
    t.om = (float)c.mg_omega;
    t.g0 = g0;
    t.f0 = c.lev[0].f32;
    t.u0 = c.lev[0].u32;
    t.M = coarse_factor();
    hipLaunchKernelGGL(k_rb_tail<CT>, dim3(1), dim3(TAIL_THREADS), 0, c.stream, t);
    DG_CHECK_LAUNCH();
    return 0;
}

template <typename CT>
int vcycle_rb(Ctx &c, int l)
{
    if (l == tail_level(c)) return launch_tail<CT>(c, l);
    Level &L = c.lev[l];
    const bool top = l == c.ng - 1;
    const int nu1 = (top || c.mg_nuc <= 0) ? c.mg_nu1 : c.mg_nuc;
    const int nu2 = (top || c.mg_nuc <= 0) ? c.mg_nu2 : c.mg_nuc;
    for (int q = 0; q < nu1; q++) {
        if (colour_pass<CT>(c, l, 0, q == 0, q == 0)) return -1;
        if (colour_pass<CT>(c, l, 1, false, q == 0)) return -1;
    }
    const RbGeom rf = rb_geom(L.g);
    const Coef<CT> cf{(const CT *)L.rb_coef, 2 * rf.half};
    Level &C = c.lev[l - 1];
    const bool csplit = l - 1 >= 1;
    const RbGeom rc = rb_geom(C.g);
    const dim3 bc = row_block(C.g.nx - 2);
    hipLaunchKernelGGL(k_rb_resid_restrict<CT>, row_grid(bc, C.g.nx - 2, C.g.ny - 2, C.g.nz - 2), bc, 0, c.stream, rf, L.rb_u, L.rb_f, cf,
                       (float)lev_h2i(c, L.g), rc, csplit ? C.rb_f : nullptr, C.g, csplit ? nullptr : C.f32);
    DG_CHECK_LAUNCH();
    if (csplit) {
        if (vcycle_rb<CT>(c, l - 1)) return -1;
    } else {
        if (coarse_solve_f32(c)) return -1;
    }
    const dim3 bf = row_block(rf.mh);
    hipLaunchKernelGGL(k_rb_prolong, row_grid(bf, rf.mh, rf.my, rf.mz), bf, 0, c.stream, rf, L.rb_u, rc,
                       csplit ? C.rb_u : nullptr, C.g, csplit ? nullptr : C.u32);
    DG_CHECK_LAUNCH();
    for (int q = 0; q < nu2; q++) {
        if (colour_pass<CT>(c, l, 1, false, 0)) return -1;
        if (colour_pass<CT>(c, l, 0, false, 0)) return -1;
    }
    return 0;
}

}  // namespace

bool rb_enabled(const Ctx &c)
{
    if (!c.rb_split || c.mg_mode != 2 || c.ng < 2 || c.mg_nu1 < 1 || c.mg_nuc < 0) return false;
    for (int l = 1; l < c.ng; l++) {
        const Grid &g = c.lev[l].g;
        if ((g.nx - 2) & 1 || (g.ny - 2) & 1 || (g.nz - 2) & 1) return false;
    }
    // the split kernels index one level with 32-bit offsets
    const Grid &gt = c.lev[c.ng - 1].g;
    if ((double)(gt.nx - 2) * (gt.ny - 2) * (gt.nz - 2) >= 2147483648.0) return false;
    return true;
}

int rb_prepare(Ctx &c)
{
    // rb_split: 1 FP16 coefficients when the densities allow, 2 FP32, 3 FP16
    const int want16 = (c.rb_split == 3 || (c.rb_split == 1 && c.coef16_ok)) ? 1 : 0;
    for (int l = 1; l < c.ng; l++) {
        Level &L = c.lev[l];
        const RbGeom r = rb_geom(L.g);
        const size_t n2 = 2 * r.half;
        if (!L.rb_u) {
            DG_HIP(hipMalloc(&L.rb_u, n2 * sizeof(float)));
            DG_HIP(hipMalloc(&L.rb_f, n2 * sizeof(float)));
            DG_HIP(hipMemset(L.rb_u, 0, n2 * sizeof(float)));
            DG_HIP(hipMemset(L.rb_f, 0, n2 * sizeof(float)));
        }
        if (L.rb_coef16 != want16) {
            if (L.rb_coef) (void)hipFree(L.rb_coef);
            L.rb_coef = nullptr;
            DG_HIP(hipMalloc(&L.rb_coef, 3 * n2 * (want16 ? sizeof(_Float16) : sizeof(float))));
            L.rb_coef16 = want16;
        }
        const dim3 bl = row_block(L.g.nx - 2), gr = row_grid(bl, L.g.nx - 2, L.g.ny - 2, L.g.nz - 2);
        if (want16)
            hipLaunchKernelGGL(k_rb_coef<_Float16>, gr, bl, 0, c.stream, L.g, r, L.a, L.b, L.c, (_Float16 *)L.rb_coef);
        else
            hipLaunchKernelGGL(k_rb_coef<float>, gr, bl, 0, c.stream, L.g, r, L.a, L.b, L.c, (float *)L.rb_coef);
        DG_CHECK_LAUNCH();
    }
    return 0;
}

int rb_load_rhs(Ctx &c, const double *r)
{
    Level &F = c.lev[c.ng - 1];
    const RbGeom rg = rb_geom(F.g);
    const dim3 bl = row_block(F.g.nx - 2), gr = row_grid(bl, F.g.nx - 2, F.g.ny - 2, F.g.nz - 2);
    hipLaunchKernelGGL(k_rb_in, gr, bl, 0, c.stream, F.g, rg, r, F.rb_f);
    DG_CHECK_LAUNCH();
    return 0;
}

int rb_vcycle(Ctx &c, const double *r)
{
    Level &F = c.lev[c.ng - 1];
    if (r && rb_load_rhs(c, r)) return -1;
    return F.rb_coef16 == 1 ? vcycle_rb<_Float16>(c, c.ng - 1) : vcycle_rb<float>(c, c.ng - 1);
}

ZSplit rb_top_z(const Ctx &c)
{
    const Level &F = c.lev[c.ng - 1];
    return ZSplit{F.rb_u, rb_geom(F.g)};
}

}  // namespace dg
