// dg_rb.hip -- the FP32 V-cycle of tolerance mode 2 on red-black split
// storage (layout in dg_rb.h).  Same algorithm as the natural-layout
// V-cycle in dg_mg.hip (symmetric red-black Gauss-Seidel V(nu1,nu2) with
// the operator's own diagonal, 8-child average restriction, injection
// prolongation, acfromac coarse operators, FP64 Cholesky on the coarsest
// level); the coefficients may be rounded to FP16.

#include <hip/hip_runtime.h>

#include "dg_prof.h"
#include "dg_rb.h"
#include "dg_util.h"

namespace dg {

int coarse_solve_f32(Ctx &c);  // dg_mg.hip

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

template <typename CT>
__global__ void k_rb_coef(Grid g, RbGeom r, const double *a, const double *b, const double *c, CT *out)
{
    const size_t n2 = 2 * r.half;
    const int mx = g.nx - 2, my = g.ny - 2, mz = g.nz - 2;
    const size_t N = (size_t)mx * my * mz;
    for (size_t p = (size_t)blockIdx.x * 256 + threadIdx.x; p < N; p += (size_t)gridDim.x * 256) {
        const int i = (int)(p % mx) + 2, j = (int)((p / mx) % my) + 2, k = (int)(p / ((size_t)mx * my)) + 2;
        const int o = g.at(i, j, k);
        const size_t e = r.at(i, j, k);
        out[e] = (CT)a[o];
        out[n2 + e] = (CT)b[o];
        out[2 * n2 + e] = (CT)c[o];
    }
}

// finest-level right-hand side: f (split FP32) = r (natural FP64)
__global__ void k_rb_in(Grid g, RbGeom r, const double *src, float *dst)
{
    const int mx = g.nx - 2, my = g.ny - 2, mz = g.nz - 2;
    const size_t N = (size_t)mx * my * mz;
    for (size_t p = (size_t)blockIdx.x * 256 + threadIdx.x; p < N; p += (size_t)gridDim.x * 256) {
        const int i = (int)(p % mx) + 2, j = (int)((p / mx) % my) + 2, k = (int)(p / ((size_t)mx * my)) + 2;
        dst[r.at(i, j, k)] = (float)src[g.at(i, j, k)];
    }
}

// One Gauss-Seidel colour pass: thread (m, j, k) updates the colour's cell
// i = 2 + ((colour+j+k)&1) + 2m of row (j,k).  Neighbours (all of the other
// colour) and the other colour's face coefficients (west, south, bottom) sit
// in the other half at m+p-1 / m+p (x) or m (y, z).  Boundary folding as the
// natural-layout pass: a missing x-min / y / z neighbour drops its face, the
// x-max neighbour is 0 with its face kept in the diagonal.  ZERO: the
// iterate is zero on entry (first pass of a V-cycle level), u is not read.
// TOP only separates the finest level's launches into their own symbol.
template <typename CT, bool ZERO, bool TOP>
__global__ void __launch_bounds__(256) k_rb_colour(RbGeom r, float *__restrict__ u, const float *__restrict__ f,
                                                   Coef<CT> cf, float h2, int colour, float om, int zown)
{
    const int m = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) + 2;
    const int k = (int)blockIdx.z + 2;
    if (m >= r.mh || j > r.ny - 1) return;
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
    const size_t e = r.at(i, j, k);
    const float u0 = u[e];
    const float ue = i < r.nx - 1 ? u[r.at(i + 1, j, k)] : 0.f;
    const float uw = i > 2 ? u[r.at(i - 1, j, k)] : u0;
    const float un = j < r.ny - 1 ? u[r.at(i, j + 1, k)] : u0;
    const float us = j > 2 ? u[r.at(i, j - 1, k)] : u0;
    const float ut = k < r.nz - 1 ? u[r.at(i, j, k + 1)] : u0;
    const float ub = k > 2 ? u[r.at(i, j, k - 1)] : u0;
    const float aw = i > 2 ? cf.a(r.at(i - 1, j, k)) : 0.f;
    const float bs = j > 2 ? cf.b(r.at(i, j - 1, k)) : 0.f;
    const float cb = k > 2 ? cf.c(r.at(i, j, k - 1)) : 0.f;
    const float s = cf.a(e) * (ue - u0) + aw * (uw - u0) + cf.b(e) * (un - u0) + bs * (us - u0) + cf.c(e) * (ut - u0) +
                    cb * (ub - u0);
    return f[e] - h2i * s;
}

// coarse right-hand side = average of the 8 children's residuals; the
// coarse level is split (fcs) or the natural-layout coarsest level (fcn)
template <typename CT>
__global__ void __launch_bounds__(256) k_rb_resid_restrict(RbGeom rf, const float *u, const float *f, Coef<CT> cf,
                                                           float h2i, RbGeom rc, float *fcs, Grid gc, float *fcn)
{
    const int mx = gc.nx - 2, my = gc.ny - 2, mz = gc.nz - 2;
    const size_t N = (size_t)mx * my * mz;
    for (size_t p = (size_t)blockIdx.x * 256 + threadIdx.x; p < N; p += (size_t)gridDim.x * 256) {
        const int I = (int)(p % mx), J = (int)((p / mx) % my), K = (int)(p / ((size_t)mx * my));
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
}

// fine u += coarse u of the parent cell (injection), fine in split order
__global__ void k_rb_prolong(RbGeom rf, float *u, RbGeom rc, const float *ucs, Grid gc, const float *ucn)
{
    const size_t n2 = 2 * rf.half;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < n2; e += (size_t)gridDim.x * 256) {
        const int colour = e >= rf.half ? 1 : 0;
        const size_t q = e - (size_t)colour * rf.half;
        const int m = (int)(q % rf.mh);
        const int j = (int)((q / rf.mh) % rf.my) + 2;
        const int k = (int)(q / ((size_t)rf.mh * rf.my)) + 2;
        const int i = 2 + ((colour + j + k) & 1) + 2 * m;
        const int ic = (i - 2) / 2 + 2, jc = (j - 2) / 2 + 2, kc = (k - 2) / 2 + 2;
        u[e] += ucs ? ucs[rc.at(ic, jc, kc)] : ucn[gc.at(ic, jc, kc)];
    }
}

unsigned gridn(size_t n)
{
    size_t b = (n + 255) / 256;
    if (b > 65535) b = 65535;
    if (b < 1) b = 1;
    return (unsigned)b;
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
    const dim3 grid((unsigned)((r.mh + BX - 1) / BX), (unsigned)((r.my + BY - 1) / BY), (unsigned)r.mz);
    const bool top = l == c.ng - 1;
    // algorithmic bytes: own f read and u write (one colour, FP32), the other
    // colour's u (not read from a zero iterate) and all coefficients
    const double cells = (double)r.half;
    const double bytes = cells * (zero ? 8.0 : (om == 1.f || zown ? 12.0 : 16.0)) + 6.0 * cells * (double)sizeof(CT);
    if (top) prof_begin(PK_RB_COLOUR, c.stream);
    if (r.mh % 4 == 0) {
        const dim3 g4((unsigned)((r.mh / 4 + BX - 1) / BX), (unsigned)((r.my + BY - 1) / BY), (unsigned)r.mz);
        if (top && zero)
            hipLaunchKernelGGL((k_rb_colour4<CT, true, true>), g4, interior_block(), 0, c.stream, r, L.rb_u, L.rb_f, cf, h2, colour, om, zown);
        else if (top)
            hipLaunchKernelGGL((k_rb_colour4<CT, false, true>), g4, interior_block(), 0, c.stream, r, L.rb_u, L.rb_f, cf, h2, colour, om, zown);
        else if (zero)
            hipLaunchKernelGGL((k_rb_colour4<CT, true, false>), g4, interior_block(), 0, c.stream, r, L.rb_u, L.rb_f, cf, h2, colour, om, zown);
        else
            hipLaunchKernelGGL((k_rb_colour4<CT, false, false>), g4, interior_block(), 0, c.stream, r, L.rb_u, L.rb_f, cf, h2, colour, om, zown);
    } else if (top && zero)
        hipLaunchKernelGGL((k_rb_colour<CT, true, true>), grid, interior_block(), 0, c.stream, r, L.rb_u, L.rb_f, cf, h2, colour, om, zown);
    else if (top)
        hipLaunchKernelGGL((k_rb_colour<CT, false, true>), grid, interior_block(), 0, c.stream, r, L.rb_u, L.rb_f, cf, h2, colour, om, zown);
    else if (zero)
        hipLaunchKernelGGL((k_rb_colour<CT, true, false>), grid, interior_block(), 0, c.stream, r, L.rb_u, L.rb_f, cf, h2, colour, om, zown);
    else
        hipLaunchKernelGGL((k_rb_colour<CT, false, false>), grid, interior_block(), 0, c.stream, r, L.rb_u, L.rb_f, cf, h2, colour, om, zown);
    if (top) prof_end(PK_RB_COLOUR, c.stream, bytes);
    DG_CHECK_LAUNCH();
    return 0;
}

template <typename CT>
int vcycle_rb(Ctx &c, int l)
{
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
    const size_t nc = (size_t)(C.g.nx - 2) * (C.g.ny - 2) * (C.g.nz - 2);
    hipLaunchKernelGGL(k_rb_resid_restrict<CT>, dim3(gridn(nc)), dim3(256), 0, c.stream, rf, L.rb_u, L.rb_f, cf,
                       (float)lev_h2i(c, L.g), rc, csplit ? C.rb_f : nullptr, C.g, csplit ? nullptr : C.f32);
    DG_CHECK_LAUNCH();
    if (csplit) {
        if (vcycle_rb<CT>(c, l - 1)) return -1;
    } else {
        if (coarse_solve_f32(c)) return -1;
    }
    hipLaunchKernelGGL(k_rb_prolong, dim3(gridn(2 * rf.half)), dim3(256), 0, c.stream, rf, L.rb_u, rc,
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
        const size_t N = (size_t)(L.g.nx - 2) * (L.g.ny - 2) * (L.g.nz - 2);
        if (want16)
            hipLaunchKernelGGL(k_rb_coef<_Float16>, dim3(gridn(N)), dim3(256), 0, c.stream, L.g, r, L.a, L.b, L.c,
                               (_Float16 *)L.rb_coef);
        else
            hipLaunchKernelGGL(k_rb_coef<float>, dim3(gridn(N)), dim3(256), 0, c.stream, L.g, r, L.a, L.b, L.c,
                               (float *)L.rb_coef);
        DG_CHECK_LAUNCH();
    }
    return 0;
}

int rb_load_rhs(Ctx &c, const double *r)
{
    Level &F = c.lev[c.ng - 1];
    const RbGeom rg = rb_geom(F.g);
    const size_t N = (size_t)(F.g.nx - 2) * (F.g.ny - 2) * (F.g.nz - 2);
    hipLaunchKernelGGL(k_rb_in, dim3(gridn(N)), dim3(256), 0, c.stream, F.g, rg, r, F.rb_f);
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
