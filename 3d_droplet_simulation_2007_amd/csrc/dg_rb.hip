// dg_rb.hip -- the tolerance-mode V-cycle on red-black split storage
// (layout in dg_rb.h): FP32 iterate in mode 2 (coefficients FP16 or FP32),
// FP64 iterate and coefficients in mode 1.  Same algorithm as the
// natural-layout V-cycle in dg_mg.hip (symmetric red-black Gauss-Seidel
// V(nu1,nu2) with the operator's own diagonal, 8-child average restriction,
// injection prolongation, acfromac coarse operators, FP64 Cholesky on the
// coarsest level).  Every kernel is templated on the coefficient type CT
// and the iterate type T; the arithmetic is T's, in the same order for
// either, so the FP64 form gives the natural-layout FP64 V-cycle's bits.

#include <hip/hip_runtime.h>

#include <type_traits>

#include "dg_coarse.h"
#include "dg_prof.h"
#include "dg_rb.h"
#include "dg_util.h"

namespace dg {

int coarse_solve_f32(Ctx &c);     // dg_mg.hip
int coarse_solve_f64(Ctx &c);     // dg_mg.hip
const double *coarse_factor(const Ctx &c);  // dg_mg.hip: Cholesky factor of the coarsest level

namespace {

// the type a stored coefficient is computed in: FP32 for FP16 and FP32
// storage, FP64 for FP64
template <typename CT>
struct CoefVal {
    using type = float;
};
template <>
struct CoefVal<double> {
    using type = double;
};

// three coefficient planes (a, b, c), each 2*half entries in split order
template <typename CT>
struct Coef {
    using V = typename CoefVal<CT>::type;
    const CT *p;
    size_t n2;  // 2 * half
    __device__ inline V a(size_t e) const { return (V)p[e]; }
    __device__ inline V b(size_t e) const { return (V)p[n2 + e]; }
    __device__ inline V c(size_t e) const { return (V)p[2 * n2 + e]; }
};

// thread (i, j, k) of the interior, one z plane per tile
template <typename CT>
__global__ void __launch_bounds__(256) k_rb_coef(Grid g, RbGeom r, const double *a, const double *b, const double *c,
                                                 CT *out)
{
    const uint3 tile = xcd_tile();
    const int i = (int)(tile.x * blockDim.x + threadIdx.x) + 2;
    const int j = (int)(tile.y * blockDim.y + threadIdx.y) + 2;
    const int k = (int)tile.z + 2;
    if (i > g.nx - 1 || j > g.ny - 1) return;
    const size_t n2 = 2 * r.half;
    const int o = g.at(i, j, k);
    const size_t e = r.at(i, j, k);
    out[e] = (CT)a[o];
    out[n2 + e] = (CT)b[o];
    out[2 * n2 + e] = (CT)c[o];
}

// finest-level right-hand side: f (split, the iterate's type) = r (natural FP64)
template <typename T>
__global__ void __launch_bounds__(256) k_rb_in(Grid g, RbGeom r, const double *src, T *dst)
{
    const uint3 tile = xcd_tile();
    const int i = (int)(tile.x * blockDim.x + threadIdx.x) + 2;
    const int j = (int)(tile.y * blockDim.y + threadIdx.y) + 2;
    const int k = (int)tile.z + 2;
    if (i > g.nx - 1 || j > g.ny - 1) return;
    dst[r.at(i, j, k)] = (T)src[g.at(i, j, k)];
}

// One Gauss-Seidel colour pass: thread (m, j, k) updates the colour's cell
// i = 2 + ((colour+j+k)&1) + 2m of row (j,k).  Neighbours (all of the other
// colour) and the other colour's face coefficients (west, south, bottom) sit
// in the other half at m+p-1 / m+p (x) or m (y, z).  Boundary folding as the
// natural-layout pass: a missing x-min / y / z neighbour drops its face, the
// x-max neighbour is 0 with its face kept in the diagonal.  ZERO: the
// iterate is zero on entry (first pass of a V-cycle level), u is not read.
// TOP only separates the finest level's launches into their own symbol.
// The new value of one cell of colour `colour` (entry m of row (j, k)):
// nb(m', dj, dk) is the other colour's value at entry m' of row (j+dj,
// k+dk), uold the cell's own value before the pass.  The colour passes
// and the tail both go through this arithmetic, so their results agree to
// the bit.  Every load is issued:
// a missing neighbour reads the cell's own entry (or its own row) instead,
// and the value is then selected away, so the loads of a cell are in flight
// together rather than each behind a branch.  Offsets are 32-bit (every
// split level has fewer than 2^31 entries).
template <typename CT, typename T, bool ZERO, class NB>
__device__ inline T rb_cell_value(const RbGeom &r, const NB &nb, const T *f, const T *uold_arr, const Coef<CT> &cf,
                                  T h2, int colour, T om, int zown, int m, int j, int k)
{
    const int p = (colour + j + k) & 1;
    const int i = 2 + p + 2 * m;
    const int rs = r.mh, ps = r.mh * r.my;  // row and plane strides of one colour
    const int row = ((k - 2) * r.my + (j - 2)) * r.mh;
    const int own = colour * (int)r.half + row + m;
    const int oth = (1 - colour) * (int)r.half + row;
    const bool hw = i > 2, he = i < r.nx - 1, hn = j < r.ny - 1, hs = j > 2, ht = k < r.nz - 1, hb = k > 2;
    const T ae = cf.a(own), bn = cf.b(own), ct = cf.c(own);
    const T awv = cf.a(hw ? oth + m + p - 1 : own);
    const T bsv = cf.b(hs ? oth - rs + m : own);
    const T cbv = cf.c(hb ? oth - ps + m : own);
    const T aw = hw ? awv : (T)0, bs = hs ? bsv : (T)0, cb = hb ? cbv : (T)0;
    T s = (T)0, d = ae;
    if (!ZERO) {
        const T ue = nb(he ? m + p : m, 0, 0), uw = nb(hw ? m + p - 1 : m, 0, 0);
        const T un = nb(m, hn ? 1 : 0, 0), us = nb(m, hs ? -1 : 0, 0);
        const T ut = nb(m, 0, ht ? 1 : 0), ub = nb(m, 0, hb ? -1 : 0);
        s = he ? s + ae * ue : s;
        s = hw ? s + aw * uw : s;
        s = hn ? s + bn * un : s;
        s = hs ? s + bs * us : s;
        s = ht ? s + ct * ut : s;
        s = hb ? s + cb * ub : s;
    }
    d = hw ? d + aw : d;
    d = hn ? d + bn : d;
    d = hs ? d + bs : d;
    d = ht ? d + ct : d;
    d = hb ? d + cb : d;
    const T gsv = (s - h2 * f[own]) / d;
    const T uold = (ZERO || zown || om == (T)1) ? (T)0 : uold_arr[own];  // Gauss-Seidel needs no old value
    return om * gsv + ((T)1 - om) * uold;
}

// in-place colour update of one cell, neighbours from the iterate itself
template <typename CT, typename T, bool ZERO>
__device__ inline void rb_cell(const RbGeom &r, T *u, const T *f, const Coef<CT> &cf, T h2, int colour, T om,
                               int zown, int m, int j, int k)
{
    const int row = ((k - 2) * r.my + (j - 2)) * r.mh;
    const T *o = u + (1 - colour) * (int)r.half + row;
    const int rs = r.mh, ps = r.mh * r.my;
    auto nb = [&](int mm, int dj, int dk) { return o[dj * rs + dk * ps + mm]; };
    u[colour * (int)r.half + row + m] = rb_cell_value<CT, T, ZERO>(r, nb, f, u, cf, h2, colour, om, zown, m, j, k);
}

template <typename CT, typename T, bool ZERO, bool TOP>
__global__ void __launch_bounds__(256) k_rb_colour(RbGeom r, T *__restrict__ u, const T *__restrict__ f, Coef<CT> cf,
                                                   T h2, int colour, T om, int zown)
{
    const uint3 tile = xcd_tile();
    const int m = (int)(tile.x * blockDim.x + threadIdx.x);
    const int j = (int)(tile.y * blockDim.y + threadIdx.y) + 2;
    const int k = (int)tile.z + 2;
    if (m >= r.mh || j > r.ny - 1) return;
    rb_cell<CT, T, ZERO>(r, u, f, cf, h2, colour, om, zown, m, j, k);
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
// (32-byte aligned for FP64)
__device__ inline void ld4(const double *p, double v[4])
{
    const double2 q0 = *reinterpret_cast<const double2 *>(p), q1 = *reinterpret_cast<const double2 *>(p + 2);
    v[0] = q0.x;
    v[1] = q0.y;
    v[2] = q1.x;
    v[3] = q1.y;
}

// four consecutive entries stored with vector stores
__device__ inline void st4(float *p, const float v[4]) { *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]); }
__device__ inline void st4(double *p, const double v[4])
{
    *reinterpret_cast<double2 *>(p) = make_double2(v[0], v[1]);
    *reinterpret_cast<double2 *>(p + 2) = make_double2(v[2], v[3]);
}

// Other-colour values around a group of four entries m0..m0+3 of row
// (j, k): row4(dj, dk) the four entries of row j+dj, plane k+dk; one(dm)
// the entry m0+dm of the group's own row (dm = -1 or 4).  NbGlobal reads
// the iterate in HBM.
template <typename T>
struct NbGlobal {
    const T *o;     // other colour, row (j, k), entry m0
    size_t rs, ps;  // row and plane strides of one colour
    __device__ inline void row4(int dj, int dk, T v[4]) const
    {
        ld4(o + (ptrdiff_t)dj * (ptrdiff_t)rs + (ptrdiff_t)dk * (ptrdiff_t)ps, v);
    }
    __device__ inline T one(int dm) const { return o[dm]; }
};

// The arithmetic of one colour update of four consecutive entries: the
// same operations in the same order as rb_cell, per cell.  own / oth: the
// group's offsets in its own and in the other colour's half (coefficients
// of the group's faces and of the other colour's west/south/bottom faces).
// As in rb_cell_value, every load is issued (a missing neighbour row reads
// the group's own row, a missing end entry the group's first one) and the
// values are selected afterwards, so the group's loads are in flight
// together instead of one round trip per boundary test.
template <typename CT, typename T, bool ZERO, class NB>
__device__ inline void rb_group4(const RbGeom &r, const NB &nb, const T *f, const T *uold_own, const Coef<CT> &cf,
                                 size_t own, size_t oth, T h2, int colour, T om, int zown, int m0, int j, int k,
                                 T out[4], const T *eo = nullptr)
{
    const int p = (colour + j + k) & 1;
    const size_t rs = (size_t)r.mh, ps = (size_t)r.mh * r.my;
    const bool hn = j < r.ny - 1, hs = j > 2, ht = k < r.nz - 1, hb = k > 2;
    const bool wfirst = m0 > 0, elast = m0 + 4 < r.mh;
    const bool need_old = !ZERO && !zown && om != (T)1;  // Gauss-Seidel needs no old value
    T ae[4], bn[4], ct[4], tw[4], bsv[4], cbv[4], fv[4], uo[4] = {(T)0, (T)0, (T)0, (T)0};
    ld4(cf.p + own, ae);
    ld4(cf.p + cf.n2 + own, bn);
    ld4(cf.p + 2 * cf.n2 + own, ct);
    ld4(cf.p + oth, tw);
    const T aw0v = (T)cf.a(wfirst ? oth - 1 : oth);
    ld4(cf.p + cf.n2 + (hs ? oth - rs : oth), bsv);
    ld4(cf.p + 2 * cf.n2 + (hb ? oth - ps : oth), cbv);
    ld4(f + own, fv);
    if (need_old) ld4(uold_own + own, uo);
    T uc[4], unv[4], usv[4], utv[4], ubv[4], u1v = (T)0;
    if (!ZERO) {
        nb.row4(0, 0, uc);
        // the entry beyond the group: east (m0+4) when p = 1, west (m0-1) when p = 0
        u1v = nb.one(p ? (elast ? 4 : 0) : (wfirst ? -1 : 0));
        nb.row4(hn ? 1 : 0, 0, unv);
        nb.row4(hs ? -1 : 0, 0, usv);
        nb.row4(0, ht ? 1 : 0, utv);
        nb.row4(0, hb ? -1 : 0, ubv);
    }
    // west faces: the other colour's a at m+p-1
    T aw[4], bs[4], cb[4];
    if (p) {
        for (int q = 0; q < 4; q++) aw[q] = tw[q];
    } else {
        aw[0] = wfirst ? aw0v : (T)0;
        aw[1] = tw[0];
        aw[2] = tw[1];
        aw[3] = tw[2];
    }
    for (int q = 0; q < 4; q++) {
        bs[q] = hs ? bsv[q] : (T)0;
        cb[q] = hb ? cbv[q] : (T)0;
    }
    if (need_old && eo)  // the coarse correction not yet added to the own colour
        for (int q = 0; q < 4; q++) uo[q] = uo[q] + eo[q];
    T ue[4] = {(T)0, (T)0, (T)0, (T)0}, uw[4] = {(T)0, (T)0, (T)0, (T)0};
    if (!ZERO) {
        if (p) {  // west m .. , east m+1 ..
            for (int q = 0; q < 4; q++) uw[q] = uc[q];
            ue[0] = uc[1];
            ue[1] = uc[2];
            ue[2] = uc[3];
            ue[3] = elast ? u1v : (T)0;
        } else {  // west m-1 .., east m ..
            for (int q = 0; q < 4; q++) ue[q] = uc[q];
            uw[0] = wfirst ? u1v : (T)0;
            uw[1] = uc[0];
            uw[2] = uc[1];
            uw[3] = uc[2];
        }
    }
    for (int q = 0; q < 4; q++) {
        const int i = 2 + p + 2 * (m0 + q);
        const bool hw = i > 2, he = i < r.nx - 1;
        T s = (T)0, d = ae[q];
        if (!ZERO) {
            s = he ? s + ae[q] * ue[q] : s;
            s = hw ? s + aw[q] * uw[q] : s;
            s = hn ? s + bn[q] * unv[q] : s;
            s = hs ? s + bs[q] * usv[q] : s;
            s = ht ? s + ct[q] * utv[q] : s;
            s = hb ? s + cb[q] * ubv[q] : s;
        }
        d = hw ? d + aw[q] : d;
        d = hn ? d + bn[q] : d;
        d = hs ? d + bs[q] : d;
        d = ht ? d + ct[q] : d;
        d = hb ? d + cb[q] : d;
        const T gsv = (s - h2 * fv[q]) / d;
        out[q] = om * gsv + ((T)1 - om) * uo[q];
    }
}

// The colour pass with four consecutive cells of a row per thread (rows of
// a multiple of 4 cells): one vector load per neighbour row instead of
// four scalar loads, same arithmetic in the same order per cell.
template <typename CT, typename T, bool ZERO, bool TOP>
__global__ void __launch_bounds__(256) k_rb_colour4(RbGeom r, T *__restrict__ u, const T *__restrict__ f, Coef<CT> cf,
                                                    T h2, int colour, T om, int zown)
{
    const uint3 tile = xcd_tile();
    const int m0 = 4 * (int)(tile.x * blockDim.x + threadIdx.x);
    const int j = (int)(tile.y * blockDim.y + threadIdx.y) + 2;
    const int k = (int)tile.z + 2;
    if (m0 >= r.mh || j > r.ny - 1) return;
    const size_t row = ((size_t)(k - 2) * r.my + (j - 2)) * r.mh;
    const size_t own = (size_t)colour * r.half + row + m0;
    const size_t oth = (size_t)(1 - colour) * r.half + row + m0;  // other colour, same row, m0
    const NbGlobal<T> nb{u + oth, (size_t)r.mh, (size_t)r.mh * r.my};
    T out[4];
    rb_group4<CT, T, ZERO>(r, nb, f, u, cf, own, oth, h2, colour, om, zown, m0, j, k, out);
    st4(u + own, out);
}

// coarse right-hand side = average of the 8 children's residuals; the
// coarse level is split (fcs) or the natural-layout coarsest level (fcn).
// The children of coarse (I, J, K) start at the even fine corner (2I+2,
// 2J+2, 2K+2), so fine cell (i0+dx, j0+dy, k0+dz) has colour (dx+dy+dz)&1
// and entry I + (dx>>1) of its row: every offset is the corner's row offset
// plus constants.  Each child's residual f - A u folds the boundary like
// the natural-layout resid (k_resid_fold), in the same operation order.
template <typename CT, typename T>
__device__ inline void rb_restrict_cell(const RbGeom &rf, const T *u, const T *f, const Coef<CT> &cf, T h2i,
                                        const RbGeom &rc, T *fcs, const Grid &gc, T *fcn, int I, int J, int K)
{
    const int i0 = 2 * I + 2, j0 = 2 * J + 2, k0 = 2 * K + 2;
    const int rs = rf.mh, ps = rf.mh * rf.my;
    const unsigned half = (unsigned)rf.half;
    const unsigned base = (unsigned)(((k0 - 2) * rf.my + (j0 - 2)) * rf.mh + I);
    auto at = [&](int dx, int dy, int dz) {
        return (unsigned)((dx + dy + dz) & 1) * half + base + (unsigned)(dy * rs + dz * ps + (dx >> 1));
    };
    // one child's residual, added to t in the children's order
    auto child = [&](T &t, int dx, int dy, int dz) {
        const int i = i0 + dx, j = j0 + dy, k = k0 + dz;
        const unsigned e = at(dx, dy, dz);
        const bool he = i < rf.nx - 1, hw = i > 2, hn = j < rf.ny - 1, hs = j > 2, ht = k < rf.nz - 1,
                   hb = k > 2;
        // every load issued, a missing neighbour reading the child
        // itself; the value is selected afterwards
        const T u0 = u[e];
        const T uev = u[he ? at(dx + 1, dy, dz) : e], uwv = u[hw ? at(dx - 1, dy, dz) : e];
        const T unv = u[hn ? at(dx, dy + 1, dz) : e], usv = u[hs ? at(dx, dy - 1, dz) : e];
        const T utv = u[ht ? at(dx, dy, dz + 1) : e], ubv = u[hb ? at(dx, dy, dz - 1) : e];
        const T awv = (T)cf.a(hw ? at(dx - 1, dy, dz) : e), bsv = (T)cf.b(hs ? at(dx, dy - 1, dz) : e),
                cbv = (T)cf.c(hb ? at(dx, dy, dz - 1) : e);
        const T ue = he ? uev : (T)0, uw = hw ? uwv : u0, un = hn ? unv : u0, us = hs ? usv : u0,
                ut = ht ? utv : u0, ub = hb ? ubv : u0;
        const T aw = hw ? awv : (T)0, bs = hs ? bsv : (T)0, cb = hb ? cbv : (T)0;
        const T s = (T)cf.a(e) * (ue - u0) + aw * (uw - u0) + (T)cf.b(e) * (un - u0) + bs * (us - u0) +
                    (T)cf.c(e) * (ut - u0) + cb * (ub - u0);
        t += f[e] - h2i * s;
    };
    T t = (T)0;
    if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int dz = 0; dz < 2; dz++)
#pragma unroll
            for (int dy = 0; dy < 2; dy++)
#pragma unroll
                for (int dx = 0; dx < 2; dx++) child(t, dx, dy, dz);
    } else {
        // FP64: four children's loads at a time (all eight in flight need
        // over 200 registers and halve the occupancy)
#pragma unroll 1
        for (int dz = 0; dz < 2; dz++)
#pragma unroll
            for (int dy = 0; dy < 2; dy++)
#pragma unroll
                for (int dx = 0; dx < 2; dx++) child(t, dx, dy, dz);
    }
    const T v = (T)0.125 * t;
    if (fcs)
        fcs[rc.at(I + 2, J + 2, K + 2)] = v;
    else
        fcn[gc.at(I + 2, J + 2, K + 2)] = v;
}

// thread (I, J, K) of the coarse interior, one z plane per tile
template <typename CT, typename T>
__global__ void __launch_bounds__(256) k_rb_resid_restrict(RbGeom rf, const T *u, const T *f, Coef<CT> cf, T h2i,
                                                           RbGeom rc, T *fcs, Grid gc, T *fcn)
{
    const uint3 tile = xcd_tile();
    const int I = (int)(tile.x * blockDim.x + threadIdx.x);
    const int J = (int)(tile.y * blockDim.y + threadIdx.y);
    const int K = (int)tile.z;
    if (I >= gc.nx - 2 || J >= gc.ny - 2) return;
    rb_restrict_cell<CT, T>(rf, u, f, cf, h2i, rc, fcs, gc, fcn, I, J, K);
}

// fine u += coarse u of the parent cell (injection), fine in split order:
// entry m of row (j, k) in both colours, cells i = 2 + p + 2m with p = 0, 1,
// both children of coarse (m+2, j', k')
template <typename T>
struct CoarseE {
    RbGeom rc;  // coarse level, split order (ucs), or
    const T *ucs;
    Grid gc;    // natural order (ucn)
    const T *ucn;
    // the correction of fine entry m of row (j, k), either colour (one load
    // from the selected array, no branch around it)
    __device__ inline T at(int m, int j, int k) const
    {
        const int ic = m + 2, jc = (j - 2) / 2 + 2, kc = (k - 2) / 2 + 2;
        const size_t e = ucs ? (size_t)rc.at32(ic, jc, kc) : (size_t)gc.at(ic, jc, kc);
        return (ucs ? ucs : ucn)[e];
    }
};

template <typename T>
__device__ inline void rb_prolong_entry(const RbGeom &rf, T *u, const RbGeom &rc, const T *ucs, const Grid &gc,
                                        const T *ucn, int m, int j, int k)
{
    const T e = CoarseE<T>{rc, ucs, gc, ucn}.at(m, j, k);
    const size_t o = ((size_t)(k - 2) * rf.my + (j - 2)) * rf.mh + m;
    u[o] += e;
    u[rf.half + o] += e;
}

template <typename T>
__global__ void __launch_bounds__(256) k_rb_prolong(RbGeom rf, T *u, RbGeom rc, const T *ucs, Grid gc, const T *ucn)
{
    const uint3 tile = xcd_tile();
    const int m = (int)(tile.x * blockDim.x + threadIdx.x);
    const int j = (int)(tile.y * blockDim.y + threadIdx.y) + 2;
    const int k = (int)tile.z + 2;
    if (m >= rf.mh || j > rf.ny - 1) return;
    rb_prolong_entry(rf, u, rc, ucs, gc, ucn, m, j, k);
}

// Other-colour neighbours with the coarse correction added as they are
// read: the values k_rb_prolong would have left in HBM.  eo: the
// corrections of entries m0 .. m0+3 of the group's own row, which a
// neighbour row with the same coarse parent row shares; ey / ez: those of
// the one y / z neighbour row with another parent (loaded up front, so a
// row read selects among registers instead of loading behind a branch).
template <typename T>
struct NbProlong {
    NbGlobal<T> g;
    CoarseE<T> e;
    int m0, j, k;
    const T *eo;
    T ey[4], ez[4];
    __device__ inline NbProlong(const NbGlobal<T> &g_, const CoarseE<T> &e_, int m0_, int j_, int k_, const T *eo_,
                                const RbGeom &r)
        : g(g_), e(e_), m0(m0_), j(j_), k(k_), eo(eo_)
    {
        // the y / z neighbour row whose parent differs (clamped to the own
        // row where it does not exist; its values are then never used)
        int jo = ((j - 2) & 1) ? j + 1 : j - 1, ko = ((k - 2) & 1) ? k + 1 : k - 1;
        jo = (jo < 2 || jo > r.ny - 1) ? j : jo;
        ko = (ko < 2 || ko > r.nz - 1) ? k : ko;
        for (int q = 0; q < 4; q++) {
            ey[q] = e.at(m0 + q, jo, k);
            ez[q] = e.at(m0 + q, j, ko);
        }
    }
    __device__ inline void row4(int dj, int dk, T v[4]) const
    {
        g.row4(dj, dk, v);
        const bool same = ((j - 2 + dj) >> 1) == ((j - 2) >> 1) && ((k - 2 + dk) >> 1) == ((k - 2) >> 1);
        for (int q = 0; q < 4; q++) v[q] = v[q] + (same ? eo[q] : (dj != 0 ? ey[q] : ez[q]));
    }
    __device__ inline T one(int dm) const { return g.one(dm) + e.at(m0 + dm, j, k); }
};

// The first post-smoothing sweep with the prolongation folded in: the pass
// of colour 1 (NBE) reads both colours as u + e (e: the coarse correction
// of the cell), the pass of colour 0 its own old values as u + e; each
// colour's corrected values are written by its own pass.  The same values
// and arithmetic as k_rb_prolong followed by the two colour passes, without
// the prolongation's read and write of both colours.
template <typename CT, typename T, bool TOP, bool NBE>
__global__ void __launch_bounds__(256) k_rb_colour4_pro(RbGeom r, T *__restrict__ u, const T *__restrict__ f,
                                                        Coef<CT> cf, T h2, int colour, T om, CoarseE<T> ce)
{
    const uint3 tile = xcd_tile();
    const int m0 = 4 * (int)(tile.x * blockDim.x + threadIdx.x);
    const int j = (int)(tile.y * blockDim.y + threadIdx.y) + 2;
    const int k = (int)tile.z + 2;
    if (m0 >= r.mh || j > r.ny - 1) return;
    const size_t row = ((size_t)(k - 2) * r.my + (j - 2)) * r.mh;
    const size_t own = (size_t)colour * r.half + row + m0;
    const size_t oth = (size_t)(1 - colour) * r.half + row + m0;
    const NbGlobal<T> nbg{u + oth, (size_t)r.mh, (size_t)r.mh * r.my};
    T eo[4], out[4];
    for (int q = 0; q < 4; q++) eo[q] = ce.at(m0 + q, j, k);
    if (NBE) {
        const NbProlong<T> nb(nbg, ce, m0, j, k, eo, r);
        rb_group4<CT, T, false>(r, nb, f, u, cf, own, oth, h2, colour, om, 0, m0, j, k, out, eo);
    } else {
        rb_group4<CT, T, false>(r, nbg, f, u, cf, own, oth, h2, colour, om, 0, m0, j, k, out, eo);
    }
    st4(u + own, out);
}

// Tile smoother of the levels between the finest and the tail: all of a
// level's pre- (or post-) smoothing sweeps in one launch.  Each workgroup
// owns a tile of cells (TsShape); it keeps the iterate on the tile
// grown by two cells on every side (the box) in LDS, runs the red-black
// sweeps on the tile grown by one cell, and writes back the tile's own
// cells.  The grown cells' neighbours outside it keep their values from
// before the sweeps (restricted additive Schwarz with an overlap of one
// cell): neighbouring tiles do not wait for each other, so the sweeps need
// no grid-wide synchronisation.  A different smoother from the global
// passes (tiles see each other's values once per launch); the CPU
// statement carries the same smoother.  Per cell the arithmetic is
// rb_cell_value's, in the same order, except that the Gauss-Seidel value
// multiplies by the reciprocal of the diagonal, formed once per launch.
//   POST = false: pre-smoothing from a zero iterate, colour 0 then 1; the
//   iterate is not read, u is written;
//   POST = true: post-smoothing, colour 1 then 0, the iterate read as
//   u + e (the coarse correction of the cell's parent), so the
//   prolongation is folded in as in k_rb_colour4_pro; out of place into
//   uout (the level's second iterate buffer), since the boxes read cells
//   that other tiles write.
// Thread t owns pairs t, t + 1024, ... of the grown tile (a pair: the cell
// of colour 0 and the cell of colour 1 at one place), whose face
// coefficients, diagonal and h^2 f it loads once into registers; a sweep
// then reads only the iterate's neighbours from LDS.
//
// The tile shapes: 16x8x8 cells on the levels of at most 2^18 cells,
// 32x16x8 above (fewer, larger tiles: one round of workgroups at 128^3, a
// smaller share of recomputed halo, and a smoother closer to the global one).
template <int X, int Y, int Z>
struct TsShape {
    static constexpr int TX = X, TY = Y, TZ = Z;
    static constexpr int BX = X + 4, BY = Y + 4, BZ = Z + 4, BN = BX * BY * BZ;  // box
    static constexpr int PX = (X + 2) / 2, GY = Y + 2, GZ = Z + 2, PAIRS = PX * GY * GZ;  // grown tile
};
using TsSmall = TsShape<16, 8, 8>;
using TsLarge = TsShape<32, 16, 8>;
constexpr int TS_THREADS = 1024;
constexpr size_t TS_MIN_CELLS = 4096;  // levels of more cells than this (at most this many: the tail)
constexpr size_t TS_MAX_CELLS_1 = (size_t)1 << 18;  // rb_tiles = 1: levels of at most this many cells
constexpr size_t TS_LARGE_CELLS = (size_t)1 << 18;  // levels of more cells than this take TsLarge

template <class S, typename T>
constexpr size_t tsmooth_lds_bytes()
{
    return (size_t)S::BN * sizeof(T);
}

// one cell of the grown tile as a thread smooths it
template <typename T>
struct TsCell {
    int o;  // LDS offset
    bool ok, he, hw, hn, hs, ht, hb;
    T ae, aw, bn, bs, ct, cb, rd, hf;  // faces (missing ones 0), 1 / diagonal, h^2 f
};

template <class S, typename CT, typename T>
__device__ inline TsCell<T> ts_cell(const RbGeom &r, const T *f, const Coef<CT> &cf, T h2, int i, int j, int k,
                                    int bi, int bj, int bk, bool own)
{
    TsCell<T> c;
    c.ok = own && i >= 2 && i <= r.nx - 1 && j >= 2 && j <= r.ny - 1 && k >= 2 && k <= r.nz - 1;
    c.o = ((k - bk) * S::BY + (j - bj)) * S::BX + (i - bi);
    c.hw = i > 2;
    c.he = i < r.nx - 1;
    c.hn = j < r.ny - 1;
    c.hs = j > 2;
    c.ht = k < r.nz - 1;
    c.hb = k > 2;
    // every load issued (a cell or face outside the level reads the nearest
    // cell inside it) and selected afterwards
    const int ic = min(max(i, 2), r.nx - 1), jc = min(max(j, 2), r.ny - 1), kc = min(max(k, 2), r.nz - 1);
    const unsigned e = r.at32(ic, jc, kc);
    const unsigned ew = r.at32(max(ic - 1, 2), jc, kc), es = r.at32(ic, max(jc - 1, 2), kc),
                   eb = r.at32(ic, jc, max(kc - 1, 2));
    const T fv = f[e];
    c.ae = (T)cf.a(e);
    c.bn = (T)cf.b(e);
    c.ct = (T)cf.c(e);
    const T awv = (T)cf.a(ew), bsv = (T)cf.b(es), cbv = (T)cf.c(eb);
    c.aw = c.hw ? awv : (T)0;
    c.bs = c.hs ? bsv : (T)0;
    c.cb = c.hb ? cbv : (T)0;
    T d = c.ae;
    d = c.hw ? d + c.aw : d;
    d = c.hn ? d + c.bn : d;
    d = c.hs ? d + c.bs : d;
    d = c.ht ? d + c.ct : d;
    d = c.hb ? d + c.cb : d;
    c.rd = (T)1 / d;  // a multiplication per sweep instead of a division
    c.hf = h2 * fv;
    return c;
}

template <class S, typename T>
__device__ inline void ts_update(T *su, const TsCell<T> &c, T om)
{
    if (!c.ok) return;
    const int o = c.o;
    T s = (T)0;
    s = c.he ? s + c.ae * su[o + 1] : s;
    s = c.hw ? s + c.aw * su[o - 1] : s;
    s = c.hn ? s + c.bn * su[o + S::BX] : s;
    s = c.hs ? s + c.bs * su[o - S::BX] : s;
    s = c.ht ? s + c.ct * su[o + S::BX * S::BY] : s;
    s = c.hb ? s + c.cb * su[o - S::BX * S::BY] : s;
    const T gsv = (s - c.hf) * c.rd;
    const T uold = om == (T)1 ? (T)0 : su[o];
    su[o] = om * gsv + ((T)1 - om) * uold;
}

template <class S, typename CT, typename T, bool POST>
__global__ void __launch_bounds__(TS_THREADS) k_rb_tsmooth(RbGeom r, const T *u, T *uout, const T *__restrict__ f,
                                                           Coef<CT> cf, T h2, T om, int nsw, CoarseE<T> ce, int ntx,
                                                           int nty)
{
    extern __shared__ __align__(16) unsigned char ts_lds[];
    T *const su = reinterpret_cast<T *>(ts_lds);
    // tiles in XCD-contiguous runs (the dispatcher deals blocks round-robin
    // to the 8 XCDs), so the boxes' shared halo cells sit in one L2
    const unsigned n = gridDim.x;
    unsigned id = blockIdx.x;
    const unsigned per = n / 8;
    if (id < per * 8) id = (id % 8) * per + id / 8;
    const int tx = (int)(id % (unsigned)ntx), ty = (int)((id / (unsigned)ntx) % (unsigned)nty),
              tz = (int)(id / ((unsigned)ntx * (unsigned)nty));
    const int i0 = 2 + tx * S::TX, j0 = 2 + ty * S::TY, k0 = 2 + tz * S::TZ;
    const int bi = i0 - 2, bj = j0 - 2, bk = k0 - 2;  // the box's first cell
    const int tid = (int)threadIdx.x;
    // the box: zero (pre-smoothing) or u + e (post-smoothing); every load
    // issued before the first store to LDS, a place outside the level
    // loading the nearest cell inside it and storing 0
    constexpr int NIT = (S::BN + TS_THREADS - 1) / TS_THREADS;
    if (POST) {
        T ru[NIT], re[NIT];
#pragma unroll
        for (int it = 0; it < NIT; it++) {
            const int q = min(tid + it * TS_THREADS, S::BN - 1);
            const int i = min(max(bi + q % S::BX, 2), r.nx - 1), j = min(max(bj + (q / S::BX) % S::BY, 2), r.ny - 1),
                      k = min(max(bk + q / (S::BX * S::BY), 2), r.nz - 1);
            ru[it] = u[r.at32(i, j, k)];
            re[it] = ce.at((i - 2) >> 1, j, k);
        }
#pragma unroll
        for (int it = 0; it < NIT; it++) {
            const int q = tid + it * TS_THREADS;
            if (q < S::BN) {
                const int i = bi + q % S::BX, j = bj + (q / S::BX) % S::BY, k = bk + q / (S::BX * S::BY);
                const bool in = i >= 2 && i <= r.nx - 1 && j >= 2 && j <= r.ny - 1 && k >= 2 && k <= r.nz - 1;
                su[q] = in ? ru[it] + re[it] : (T)0;
            }
        }
    } else {
        for (int q = tid; q < S::BN; q += TS_THREADS) su[q] = (T)0;
    }
    // this thread's pairs of the grown tile (pair tid + p TS_THREADS), both
    // colours
    constexpr int PP = (S::PAIRS + TS_THREADS - 1) / TS_THREADS;
    TsCell<T> c0[PP], c1[PP];
    const int gi0 = i0 - 1;
#pragma unroll
    for (int pp = 0; pp < PP; pp++) {
        const int e = tid + pp * TS_THREADS;
        const bool own = e < S::PAIRS;
        const int pq = e % S::PX, j = j0 - 1 + (e / S::PX) % S::GY, k = k0 - 1 + e / (S::PX * S::GY);
        const int p0 = (gi0 + j + k) & 1;  // colour 0's cell is the pair's first when (i + j + k) is even there
        c0[pp] = ts_cell<S, CT, T>(r, f, cf, h2, gi0 + 2 * pq + p0, j, k, bi, bj, bk, own);
        c1[pp] = ts_cell<S, CT, T>(r, f, cf, h2, gi0 + 2 * pq + (1 - p0), j, k, bi, bj, bk, own);
    }
    __syncthreads();
    for (int sw = 0; sw < nsw; sw++) {
#pragma unroll
        for (int pp = 0; pp < PP; pp++) ts_update<S>(su, POST ? c1[pp] : c0[pp], om);
        __syncthreads();
#pragma unroll
        for (int pp = 0; pp < PP; pp++) ts_update<S>(su, POST ? c0[pp] : c1[pp], om);
        __syncthreads();
    }
    constexpr int CPT = S::TX * S::TY * S::TZ / TS_THREADS;
    static_assert(CPT * TS_THREADS == S::TX * S::TY * S::TZ, "whole tile cells per thread");
#pragma unroll
    for (int q0 = 0; q0 < CPT; q0++) {
        const int q = tid + q0 * TS_THREADS;
        const int i = i0 + q % S::TX, jj = j0 + (q / S::TX) % S::TY, kk = k0 + q / (S::TX * S::TY);
        if (i <= r.nx - 1 && jj <= r.ny - 1 && kk <= r.nz - 1)
            uout[r.at32(i, jj, kk)] = su[((kk - bk) * S::BY + (jj - bj)) * S::BX + (i - bi)];
    }
}

// The V-cycle below a small level in one workgroup: every colour pass,
// restriction, the coarsest solve and every prolongation of the levels
// 1..top as loops of TAIL_THREADS threads separated by barriers, with the
// per-cell arithmetic of the kernels above (same results, ~4 launches per
// level fewer).  Levels are indexed as in the context (0 = coarsest, the
// natural-layout direct solve).  LDS form: the levels' u, f and face
// coefficients and the coarsest level's f and u are staged in LDS (loaded
// once per launch: the top level's f and every coefficient; the top level's
// u written back at the end), or only u and f when the coefficients do not
// fit (FP64); the global form works in place.
constexpr int TAIL_THREADS = 1024;
constexpr int TAIL_MAX_LEVELS = 8;
constexpr size_t TAIL_MAX_CELLS = 4096;  // largest level taken into the tail
static_assert(TS_MIN_CELLS == TAIL_MAX_CELLS, "the tile smoother serves the levels above the tail");

template <typename CT, typename T>
struct RbTail {
    int top;  // highest level in the tail (>= 1)
    RbGeom r[TAIL_MAX_LEVELS];
    T *u[TAIL_MAX_LEVELS], *f[TAIL_MAX_LEVELS];
    Coef<CT> cf[TAIL_MAX_LEVELS];
    T h2[TAIL_MAX_LEVELS], h2i[TAIL_MAX_LEVELS];
    int nu1[TAIL_MAX_LEVELS], nu2[TAIL_MAX_LEVELS];
    T om;
    Grid g0;  // coarsest level (natural layout)
    T *f0, *u0;
    const double *M;
    // LDS form: byte offsets of level l's u, f and coefficients (index l),
    // and of the coarsest level's f and u (index 0: f0, index top + 1: u0)
    unsigned lu[TAIL_MAX_LEVELS + 1], lf[TAIL_MAX_LEVELS + 1], lc[TAIL_MAX_LEVELS];
    // 1: the LDS form keeps u and f in LDS but reads the coefficients where
    // they lie (FP64 coefficients of a 16^3 level exceed the LDS); the
    // kernel is instantiated for either case (its CG parameter)
    int cglob;
};

// LDS layout of the staged tail: per level u, f (type T) and 3 coefficient
// planes, then the coarsest level's f and u, 16-byte aligned pieces.
// Fills the offsets of t and returns the total.
template <typename CT, typename T>
inline size_t tail_lds_layout(RbTail<CT, T> &t)
{
    auto al = [](size_t b) { return (b + 15) & ~(size_t)15; };
    size_t b = 0;
    for (int l = 1; l <= t.top; l++) {
        const size_t n2 = 2 * t.r[l].half;
        t.lu[l] = (unsigned)b;
        b += al(n2 * sizeof(T));
        t.lf[l] = (unsigned)b;
        b += al(n2 * sizeof(T));
        t.lc[l] = (unsigned)b;
        if (!t.cglob) b += al(3 * n2 * sizeof(CT));
    }
    const size_t n0 = (size_t)t.g0.nx * t.g0.ny * t.g0.nz;
    t.lf[0] = (unsigned)b;
    b += al(n0 * sizeof(T));
    t.lu[t.top + 1] = (unsigned)b;
    b += al(n0 * sizeof(T));
    return b;
}

// one level of the tail as a pass works on it: geometry, arrays (in LDS or
// HBM) and constants, read from the kernel arguments once per visit
template <typename CT, typename T>
struct TailLev {
    RbGeom r;
    T *u, *f;
    Coef<CT> cf;
    T h2, h2i;
};

// where level l's arrays live for one form of the tail (CG: coefficients
// in HBM even in the LDS form)
template <typename CT, typename T, bool LDS, bool CG = false>
struct TailMem {
    const RbTail<CT, T> &t;
    unsigned char *lds;
    __device__ inline T *u(int l) const { return LDS ? (T *)(lds + t.lu[l]) : t.u[l]; }
    __device__ inline T *f(int l) const { return LDS ? (T *)(lds + t.lf[l]) : t.f[l]; }
    __device__ inline Coef<CT> cf(int l) const
    {
        return (LDS && !CG) ? Coef<CT>{(const CT *)(lds + t.lc[l]), 2 * t.r[l].half} : t.cf[l];
    }
    __device__ inline T *f0() const { return LDS ? (T *)(lds + t.lf[0]) : t.f0; }
    __device__ inline T *u0() const { return LDS ? (T *)(lds + t.lu[t.top + 1]) : t.u0; }
    __device__ inline TailLev<CT, T> level(int l) const
    {
        return TailLev<CT, T>{t.r[l], u(l), f(l), cf(l), t.h2[l], t.h2i[l]};
    }
};

// this thread's entries (m, j, k) of one colour of a tail level, decoded
// once per visit of the level instead of once per pass
constexpr int TAIL_CPT = (int)((TAIL_MAX_CELLS / 2 + TAIL_THREADS - 1) / TAIL_THREADS);
struct TailCells {
    int n;
    int m[TAIL_CPT], j[TAIL_CPT], k[TAIL_CPT];
};

__device__ inline TailCells tail_cells(const RbGeom &r)
{
    TailCells c;
    c.n = 0;
#pragma unroll
    for (int s = 0; s < TAIL_CPT; s++) {
        const int e = (int)threadIdx.x + s * TAIL_THREADS;
        c.m[s] = c.j[s] = c.k[s] = 0;
        if (e < (int)r.half) {
            const int q = e / r.mh;
            c.m[s] = e - q * r.mh;
            c.j[s] = q % r.my + 2;
            c.k[s] = q / r.my + 2;
            c.n = s + 1;
        }
    }
    return c;
}

template <typename CT, typename T, bool ZERO>
__device__ inline void tail_pass(const TailLev<CT, T> &L, const TailCells &cells, T om, int colour, int zown)
{
#pragma unroll
    for (int s = 0; s < TAIL_CPT; s++)
        if (s < cells.n)
            rb_cell<CT, T, ZERO>(L.r, L.u, L.f, L.cf, L.h2, colour, om, zown, cells.m[s], cells.j[s], cells.k[s]);
    __syncthreads();
}

template <typename CT, typename T, bool LDS, bool CG>
__device__ inline void tail_body(const TailMem<CT, T, LDS, CG> &mem, double *y)
{
    const RbTail<CT, T> &t = mem.t;
    const T om = t.om;
    for (int l = t.top; l >= 1; l--) {
        const TailLev<CT, T> L = mem.level(l);
        const TailCells cells = tail_cells(L.r);
        const int nu1 = t.nu1[l];
        for (int q = 0; q < nu1; q++) {
            if (q == 0)
                tail_pass<CT, T, true>(L, cells, om, 0, 1);
            else
                tail_pass<CT, T, false>(L, cells, om, 0, 0);
            tail_pass<CT, T, false>(L, cells, om, 1, q == 0);
        }
        const RbGeom &rf = L.r;
        const bool csplit = l - 1 >= 1;
        T *fc = csplit ? mem.f(l - 1) : mem.f0();
        const RbGeom rc = t.r[csplit ? l - 1 : l];
        const Grid g0 = t.g0;
        const int cx = rf.mh, cy = rf.my / 2, n = cx * cy * (rf.mz / 2);  // coarse interior
        for (int e = (int)threadIdx.x; e < n; e += TAIL_THREADS) {
            const int I = e % cx, q = e / cx;
            rb_restrict_cell<CT, T>(rf, L.u, L.f, L.cf, L.h2i, rc, csplit ? fc : nullptr, g0, csplit ? nullptr : fc, I,
                                    q % cy, q / cy);
        }
        __syncthreads();
    }
    coarse_subst<T>(t.g0, t.M, mem.f0(), mem.u0(), y, (int)threadIdx.x, TAIL_THREADS);
    __syncthreads();
    for (int l = 1; l <= t.top; l++) {
        const TailLev<CT, T> L = mem.level(l);
        const RbGeom &rf = L.r;
        const bool csplit = l - 1 >= 1;
        const T *uc = csplit ? mem.u(l - 1) : mem.u0();
        const RbGeom rc = t.r[csplit ? l - 1 : l];
        const Grid g0 = t.g0;
        const int n = (int)rf.half;
        for (int e = (int)threadIdx.x; e < n; e += TAIL_THREADS) {
            const int m = e % rf.mh, q = e / rf.mh;
            rb_prolong_entry<T>(rf, L.u, rc, csplit ? uc : nullptr, g0, csplit ? nullptr : uc, m, q % rf.my + 2,
                                q / rf.my + 2);
        }
        __syncthreads();
        const TailCells cells = tail_cells(rf);
        const int nu2 = t.nu2[l];
        for (int q = 0; q < nu2; q++) {
            tail_pass<CT, T, false>(L, cells, om, 1, 0);
            tail_pass<CT, T, false>(L, cells, om, 0, 0);
        }
    }
}

template <typename CT, typename T>
__global__ void __launch_bounds__(TAIL_THREADS) k_rb_tail(RbTail<CT, T> t)
{
    __shared__ double y[512];
    const TailMem<CT, T, false> mem{t, nullptr};
    tail_body<CT, T, false, false>(mem, y);
}

template <typename CT, typename T, bool CG>
__global__ void __launch_bounds__(TAIL_THREADS) k_rb_tail_lds(RbTail<CT, T> t)
{
    extern __shared__ __align__(16) unsigned char tail_lds[];
    __shared__ double y[512];
    const TailMem<CT, T, true, CG> mem{t, tail_lds};
    // stage the top level's right-hand side and every level's coefficients
    {
        const int l = t.top;
        const int n2 = (int)(2 * t.r[l].half);
        T *fd = mem.f(l);
        for (int e = (int)threadIdx.x; e < n2; e += TAIL_THREADS) fd[e] = t.f[l][e];
    }
    for (int l = 1; l <= t.top && !CG; l++) {
        const int n = (int)(6 * t.r[l].half);
        CT *cd = (CT *)mem.cf(l).p;
        for (int e = (int)threadIdx.x; e < n; e += TAIL_THREADS) cd[e] = t.cf[l].p[e];
    }
    __syncthreads();
    tail_body<CT, T, true, CG>(mem, y);
    {
        const int l = t.top;
        const int n2 = (int)(2 * t.r[l].half);
        const T *us = mem.u(l);
        for (int e = (int)threadIdx.x; e < n2; e += TAIL_THREADS) t.u[l][e] = us[e];
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
// (the second pass of the first sweep).  Adds the pass's algorithmic bytes
// to *bytes_sum.
template <typename CT, typename T>
int colour_pass(Ctx &c, int l, int colour, bool zero, int zown, double *bytes_sum)
{
    const T om = (T)c.mg_omega;
    Level &L = c.lev[l];
    T *u = rb_u_of<T>(L), *f = rb_f_of<T>(L);
    const RbGeom r = rb_geom(L.g);
    const Coef<CT> cf{(const CT *)L.rb_coef, 2 * r.half};
    const T h2 = (T)(1.0 / lev_h2i(c, L.g));
    const dim3 b1 = row_block(r.mh), grid = row_grid(b1, r.mh, r.my, r.mz);
    const bool top = l == c.ng - 1;
    // algorithmic bytes: own f read and u write (one colour), the other
    // colour's u (not read from a zero iterate), the own old u (SOR) and
    // all coefficients
    const double cells = (double)r.half;
    *bytes_sum += cells * (double)sizeof(T) * (zero ? 2.0 : (om == (T)1 || zown ? 3.0 : 4.0)) +
                  6.0 * cells * (double)sizeof(CT);
    if (r.mh % 4 == 0) {
        const dim3 b4 = row_block(r.mh / 4), g4 = row_grid(b4, r.mh / 4, r.my, r.mz);
        if (top && zero)
            hipLaunchKernelGGL((k_rb_colour4<CT, T, true, true>), g4, b4, 0, c.stream, r, u, f, cf, h2, colour, om, zown);
        else if (top)
            hipLaunchKernelGGL((k_rb_colour4<CT, T, false, true>), g4, b4, 0, c.stream, r, u, f, cf, h2, colour, om, zown);
        else if (zero)
            hipLaunchKernelGGL((k_rb_colour4<CT, T, true, false>), g4, b4, 0, c.stream, r, u, f, cf, h2, colour, om, zown);
        else
            hipLaunchKernelGGL((k_rb_colour4<CT, T, false, false>), g4, b4, 0, c.stream, r, u, f, cf, h2, colour, om, zown);
    } else if (top && zero)
        hipLaunchKernelGGL((k_rb_colour<CT, T, true, true>), grid, b1, 0, c.stream, r, u, f, cf, h2, colour, om, zown);
    else if (top)
        hipLaunchKernelGGL((k_rb_colour<CT, T, false, true>), grid, b1, 0, c.stream, r, u, f, cf, h2, colour, om, zown);
    else if (zero)
        hipLaunchKernelGGL((k_rb_colour<CT, T, true, false>), grid, b1, 0, c.stream, r, u, f, cf, h2, colour, om, zown);
    else
        hipLaunchKernelGGL((k_rb_colour<CT, T, false, false>), grid, b1, 0, c.stream, r, u, f, cf, h2, colour, om, zown);
    DG_CHECK_LAUNCH();
    return 0;
}

// The first post-smoothing sweep of level l (colour 1, then colour 0) with
// the prolongation from the level below folded in (rows of a multiple of 4
// entries)
template <typename CT, typename T>
int colour_pass_pro(Ctx &c, int l, const CoarseE<T> &ce, double *bytes_sum)
{
    const T om = (T)c.mg_omega;
    Level &L = c.lev[l];
    T *u = rb_u_of<T>(L), *f = rb_f_of<T>(L);
    const RbGeom r = rb_geom(L.g);
    const Coef<CT> cf{(const CT *)L.rb_coef, 2 * r.half};
    const T h2 = (T)(1.0 / lev_h2i(c, L.g));
    const bool top = l == c.ng - 1;
    // two colour passes' bytes, each also reading the coarse level (a
    // quarter as many cells as one fine colour)
    const double cells = (double)r.half, ts = (double)sizeof(T);
    *bytes_sum += 2.0 * (cells * ts * (om == (T)1 ? 3.0 : 4.0) + 6.0 * cells * (double)sizeof(CT) + cells * ts / 4.0);
    const dim3 b4 = row_block(r.mh / 4), g4 = row_grid(b4, r.mh / 4, r.my, r.mz);
    if (top) {
        hipLaunchKernelGGL((k_rb_colour4_pro<CT, T, true, true>), g4, b4, 0, c.stream, r, u, f, cf, h2, 1, om, ce);
        hipLaunchKernelGGL((k_rb_colour4_pro<CT, T, true, false>), g4, b4, 0, c.stream, r, u, f, cf, h2, 0, om, ce);
    } else {
        hipLaunchKernelGGL((k_rb_colour4_pro<CT, T, false, true>), g4, b4, 0, c.stream, r, u, f, cf, h2, 1, om, ce);
        hipLaunchKernelGGL((k_rb_colour4_pro<CT, T, false, false>), g4, b4, 0, c.stream, r, u, f, cf, h2, 0, om, ce);
    }
    DG_CHECK_LAUNCH();
    return 0;
}

// the coarsest level's natural-layout right-hand side and iterate of type T
template <typename T>
void coarsest_arrays(Level &L0, T *&f0, T *&u0);
template <>
void coarsest_arrays<float>(Level &L0, float *&f0, float *&u0)
{
    f0 = L0.f32;
    u0 = L0.u32;
}
template <>
void coarsest_arrays<double>(Level &L0, double *&f0, double *&u0)
{
    f0 = L0.f64;
    u0 = L0.u64;
}

// highest level taken into the single-workgroup tail, or -1
int tail_level(const Ctx &c)
{
    int lt = -1;
    for (int l = 1; l < c.ng - 1 && l < TAIL_MAX_LEVELS; l++)
        if (2 * rb_geom(c.lev[l].g).half <= TAIL_MAX_CELLS) lt = l;
    return lt;
}

template <typename CT, typename T>
int launch_tail(Ctx &c, int lt)
{
    const Grid &g0 = c.lev[0].g;
    if ((g0.nx - 2) * (g0.ny - 2) * (g0.nz - 2) > 512) {
        set_error("split V-cycle tail: coarsest level above 512 unknowns");
        return -1;
    }
    RbTail<CT, T> t{};
    t.top = lt;
    for (int l = 1; l <= lt; l++) {
        Level &L = c.lev[l];
        t.r[l] = rb_geom(L.g);
        t.u[l] = rb_u_of<T>(L);
        t.f[l] = rb_f_of<T>(L);
        t.cf[l] = Coef<CT>{(const CT *)L.rb_coef, 2 * t.r[l].half};
        t.h2[l] = (T)(1.0 / lev_h2i(c, L.g));
        t.h2i[l] = (T)lev_h2i(c, L.g);
        t.nu1[l] = c.mg_nuc <= 0 ? c.mg_nu1 : c.mg_nuc;
        t.nu2[l] = c.mg_nuc <= 0 ? c.mg_nu2 : c.mg_nuc;
    }
    t.om = (T)c.mg_omega;
    t.g0 = g0;
    coarsest_arrays<T>(c.lev[0], t.f0, t.u0);
    t.M = coarse_factor(c);
    // LDS form when the device's per-workgroup LDS holds the staged levels
    // next to the 4 KB substitution buffer
    static int lds_max = -1;
    if (lds_max < 0) {
        int dev = 0;
        DG_HIP(hipGetDevice(&dev));
        DG_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    }
    size_t lds = tail_lds_layout(t);
    if (lds + 512 * sizeof(double) > (size_t)lds_max) {
        // u and f in LDS, coefficients read where they lie
        t.cglob = 1;
        lds = tail_lds_layout(t);
    }
    if (c.rb_tail == 2 && lds + 512 * sizeof(double) <= (size_t)lds_max) {
        if (t.cglob) {
            if (lds > 48 * 1024)
                DG_HIP(hipFuncSetAttribute((const void *)k_rb_tail_lds<CT, T, true>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            hipLaunchKernelGGL((k_rb_tail_lds<CT, T, true>), dim3(1), dim3(TAIL_THREADS), lds, c.stream, t);
        } else {
            if (lds > 48 * 1024)
                DG_HIP(hipFuncSetAttribute((const void *)k_rb_tail_lds<CT, T, false>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            hipLaunchKernelGGL((k_rb_tail_lds<CT, T, false>), dim3(1), dim3(TAIL_THREADS), lds, c.stream, t);
        }
    } else {
        hipLaunchKernelGGL((k_rb_tail<CT, T>), dim3(1), dim3(TAIL_THREADS), 0, c.stream, t);
    }
    DG_CHECK_LAUNCH();
    return 0;
}

// whether level l's sweeps run as the tile smoother: on request, below the
// finest level, above the tail's size (rb_tiles 1: and at most 2^18 cells,
// where the global passes are launch-bound), and when its tile's box fits
// the LDS
template <typename CT, typename T>
bool tile_level(const Ctx &c, int l)
{
    if (!c.rb_tiles || l >= c.ng - 1 || l < 1) return false;
    const size_t cells = 2 * rb_geom(c.lev[l].g).half;
    if (cells <= TS_MIN_CELLS || (c.rb_tiles == 1 && cells > TS_MAX_CELLS_1)) return false;
    static int lds_max = -1;
    if (lds_max < 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            lds_max = 0;
    }
    const size_t lds = cells > TS_LARGE_CELLS ? tsmooth_lds_bytes<TsLarge, T>() : tsmooth_lds_bytes<TsSmall, T>();
    return lds <= (size_t)lds_max;
}

template <class S, typename CT, typename T, bool POST>
int tile_smooth_s(Ctx &c, int l, int nsw, const CoarseE<T> &ce)
{
    Level &L = c.lev[l];
    const RbGeom r = rb_geom(L.g);
    const Coef<CT> cf{(const CT *)L.rb_coef, 2 * r.half};
    const int ntx = (2 * r.mh + S::TX - 1) / S::TX, nty = (r.my + S::TY - 1) / S::TY, ntz = (r.mz + S::TZ - 1) / S::TZ;
    const size_t lds = tsmooth_lds_bytes<S, T>();
    if (lds > 48 * 1024)
        DG_HIP(hipFuncSetAttribute((const void *)k_rb_tsmooth<S, CT, T, POST>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_rb_tsmooth<S, CT, T, POST>), dim3((unsigned)(ntx * nty * ntz)), dim3(TS_THREADS), lds,
                       c.stream, r, rb_u_of<T>(L), POST ? rb_w_of<T>(L) : rb_u_of<T>(L), rb_f_of<T>(L), cf,
                       (T)(1.0 / lev_h2i(c, L.g)), (T)c.mg_omega, nsw, ce, ntx, nty);
    DG_CHECK_LAUNCH();
    return 0;
}

// all nsw pre- (POST false) or post-smoothing sweeps of level l in one launch
template <typename CT, typename T, bool POST>
int tile_smooth(Ctx &c, int l, int nsw, const CoarseE<T> &ce)
{
    if (2 * rb_geom(c.lev[l].g).half > TS_LARGE_CELLS) return tile_smooth_s<TsLarge, CT, T, POST>(c, l, nsw, ce);
    return tile_smooth_s<TsSmall, CT, T, POST>(c, l, nsw, ce);
}

// the coarsest level's solve in the iterate's type
template <typename T>
int coarse_solve_t(Ctx &c);
template <>
int coarse_solve_t<float>(Ctx &c) { return coarse_solve_f32(c); }
template <>
int coarse_solve_t<double>(Ctx &c) { return coarse_solve_f64(c); }

template <typename CT, typename T>
int vcycle_rb(Ctx &c, int l)
{
    if (c.rb_tail > 0 && l == tail_level(c)) return launch_tail<CT, T>(c, l);
    Level &L = c.lev[l];
    T *const lu = rb_u_of<T>(L), *const lf = rb_f_of<T>(L);
    const bool top = l == c.ng - 1;
    const int nu1 = (top || c.mg_nuc <= 0) ? c.mg_nu1 : c.mg_nuc;
    const int nu2 = (top || c.mg_nuc <= 0) ? c.mg_nu2 : c.mg_nuc;
    const bool tiles = tile_level<CT, T>(c, l);
    const int nut1 = c.rb_tile_nu > 0 ? c.rb_tile_nu : nu1, nut2 = c.rb_tile_nu > 0 ? c.rb_tile_nu : nu2;
    if (tiles) {
        if (tile_smooth<CT, T, false>(c, l, nut1, CoarseE<T>{})) return -1;
    } else {
        // the finest level's pre-smoothing passes run back to back: one
        // timing pair for the run when the family is profiled
        double pb = 0.0;
        if (top) prof_begin(PK_RB_COLOUR, c.stream);
        for (int q = 0; q < nu1; q++) {
            if (colour_pass<CT, T>(c, l, 0, q == 0, q == 0, &pb)) return -1;
            if (colour_pass<CT, T>(c, l, 1, false, q == 0, &pb)) return -1;
        }
        if (top) prof_end(PK_RB_COLOUR, c.stream, pb, 2 * nu1);
    }
    const RbGeom rf = rb_geom(L.g);
    const Coef<CT> cf{(const CT *)L.rb_coef, 2 * rf.half};
    Level &C = c.lev[l - 1];
    const bool csplit = l - 1 >= 1;
    const RbGeom rc = rb_geom(C.g);
    T *cf0 = nullptr, *cu0 = nullptr;
    coarsest_arrays<T>(C, cf0, cu0);
    // the coarse level's result: rb_u, or rb_w where the tile smoother's
    // post-smoothing wrote it
    T *const cfs = csplit ? rb_f_of<T>(C) : nullptr;
    T *const cus = csplit ? (tile_level<CT, T>(c, l - 1) ? rb_w_of<T>(C) : rb_u_of<T>(C)) : nullptr;
    const dim3 bc = row_block(C.g.nx - 2);
    hipLaunchKernelGGL((k_rb_resid_restrict<CT, T>), row_grid(bc, C.g.nx - 2, C.g.ny - 2, C.g.nz - 2), bc, 0, c.stream,
                       rf, lu, lf, cf, (T)lev_h2i(c, L.g), rc, cfs, C.g, csplit ? nullptr : cf0);
    DG_CHECK_LAUNCH();
    if (csplit) {
        if (vcycle_rb<CT, T>(c, l - 1)) return -1;
    } else {
        if (coarse_solve_t<T>(c)) return -1;
    }
    const CoarseE<T> ce{rc, cus, C.g, csplit ? nullptr : cu0};
    // the colour passes take the prolongation into their first sweep
    if (tiles) return tile_smooth<CT, T, true>(c, l, nut2, ce);
    const bool pro_folded = nu2 >= 1 && rf.mh % 4 == 0;
    if (!pro_folded) {
        const dim3 bf = row_block(rf.mh);
        hipLaunchKernelGGL(k_rb_prolong<T>, row_grid(bf, rf.mh, rf.my, rf.mz), bf, 0, c.stream, rf, lu, rc, cus, C.g,
                           csplit ? nullptr : cu0);
        DG_CHECK_LAUNCH();
    }
    {
        double pb = 0.0;
        if (top) prof_begin(PK_RB_COLOUR, c.stream);
        for (int q = 0; q < nu2; q++) {
            if (q == 0 && pro_folded) {
                if (colour_pass_pro<CT, T>(c, l, ce, &pb)) return -1;
                continue;
            }
            if (colour_pass<CT, T>(c, l, 1, false, 0, &pb)) return -1;
            if (colour_pass<CT, T>(c, l, 0, false, 0, &pb)) return -1;
        }
        if (top) prof_end(PK_RB_COLOUR, c.stream, pb, 2 * nu2);
    }
    return 0;
}

}  // namespace

bool rb_enabled(const Ctx &c)
{
    if (!c.rb_split || (c.mg_mode != 1 && c.mg_mode != 2) || c.ng < 2 || c.mg_nu1 < 1 || c.mg_nuc < 0) return false;
    for (int l = 1; l < c.ng; l++) {
        const Grid &g = c.lev[l].g;
        if ((g.nx - 2) & 1 || (g.ny - 2) & 1 || (g.nz - 2) & 1) return false;
    }
    // the split kernels index one level with 32-bit offsets
    const Grid &gt = c.lev[c.ng - 1].g;
    if ((double)(gt.nx - 2) * (gt.ny - 2) * (gt.nz - 2) >= 2147483648.0) return false;
    return true;
}

// mode 1: FP64; mode 2: rb_split 1 FP16 when the densities allow, 2 FP32, 3 FP16
int rb_coef_type(const Ctx &c)
{
    if (c.mg_mode == 1) return RB_CT_F64;
    return (c.rb_split == 3 || (c.rb_split == 1 && c.coef16_ok)) ? RB_CT_F16 : RB_CT_F32;
}

static size_t coef_bytes(int ct) { return ct == RB_CT_F16 ? 2 : (ct == RB_CT_F32 ? 4 : 8); }

int rb_alloc(Ctx &c)
{
    const int want = rb_coef_type(c);
    const bool f64 = c.mg_mode == 1;
    for (int l = 1; l < c.ng; l++) {
        Level &L = c.lev[l];
        const RbGeom r = rb_geom(L.g);
        const size_t n2 = 2 * r.half;
        if (f64 && !L.rb_u64) {
            DG_HIP(hipMalloc(&L.rb_u64, n2 * sizeof(double)));
            DG_HIP(hipMalloc(&L.rb_f64, n2 * sizeof(double)));
            DG_HIP(hipMemset(L.rb_u64, 0, n2 * sizeof(double)));
            DG_HIP(hipMemset(L.rb_f64, 0, n2 * sizeof(double)));
        }
        if (!f64 && !L.rb_u) {
            DG_HIP(hipMalloc(&L.rb_u, n2 * sizeof(float)));
            DG_HIP(hipMalloc(&L.rb_f, n2 * sizeof(float)));
            DG_HIP(hipMemset(L.rb_u, 0, n2 * sizeof(float)));
            DG_HIP(hipMemset(L.rb_f, 0, n2 * sizeof(float)));
        }
        if (!f64 && (c.rb_tiles && l < c.ng - 1) && !L.rb_w) {
            DG_HIP(hipMalloc(&L.rb_w, n2 * sizeof(float)));
            DG_HIP(hipMemset(L.rb_w, 0, n2 * sizeof(float)));
        }
        if (f64 && c.rb_tiles && l < c.ng - 1 && !L.rb_w64) {
            DG_HIP(hipMalloc(&L.rb_w64, n2 * sizeof(double)));
            DG_HIP(hipMemset(L.rb_w64, 0, n2 * sizeof(double)));
        }
        if (L.rb_ct != want) {
            if (L.rb_coef) (void)hipFree(L.rb_coef);
            L.rb_coef = nullptr;
            DG_HIP(hipMalloc(&L.rb_coef, 3 * n2 * coef_bytes(want)));
            L.rb_ct = want;
        }
    }
    return 0;
}

int rb_prepare(Ctx &c)
{
    if (rb_alloc(c)) return -1;
    const int want = rb_coef_type(c);
    for (int l = 1; l < c.ng; l++) {
        Level &L = c.lev[l];
        const RbGeom r = rb_geom(L.g);
        if (l == c.ng - 1 && c.top_coef_ready) continue;  // written by solvelpres's abcei
        if (l < c.ng - 1 && c.coarse_coef_ready) continue;  // written by acfromac
        const dim3 bl = row_block(L.g.nx - 2), gr = row_grid(bl, L.g.nx - 2, L.g.ny - 2, L.g.nz - 2);
        if (want == RB_CT_F16)
            hipLaunchKernelGGL(k_rb_coef<_Float16>, gr, bl, 0, c.stream, L.g, r, L.a, L.b, L.c, (_Float16 *)L.rb_coef);
        else if (want == RB_CT_F32)
            hipLaunchKernelGGL(k_rb_coef<float>, gr, bl, 0, c.stream, L.g, r, L.a, L.b, L.c, (float *)L.rb_coef);
        else
            hipLaunchKernelGGL(k_rb_coef<double>, gr, bl, 0, c.stream, L.g, r, L.a, L.b, L.c, (double *)L.rb_coef);
        DG_CHECK_LAUNCH();
    }
    c.top_coef_ready = 0;
    c.coarse_coef_ready = 0;
    return 0;
}

int rb_load_rhs(Ctx &c, const double *r)
{
    Level &F = c.lev[c.ng - 1];
    const RbGeom rg = rb_geom(F.g);
    const dim3 bl = row_block(F.g.nx - 2), gr = row_grid(bl, F.g.nx - 2, F.g.ny - 2, F.g.nz - 2);
    if (c.mg_mode == 1)
        hipLaunchKernelGGL(k_rb_in<double>, gr, bl, 0, c.stream, F.g, rg, r, F.rb_f64);
    else
        hipLaunchKernelGGL(k_rb_in<float>, gr, bl, 0, c.stream, F.g, rg, r, F.rb_f);
    DG_CHECK_LAUNCH();
    return 0;
}

int rb_vcycle(Ctx &c, const double *r)
{
    Level &F = c.lev[c.ng - 1];
    if (r && rb_load_rhs(c, r)) return -1;
    switch (F.rb_ct) {
    case RB_CT_F16: return vcycle_rb<_Float16, float>(c, c.ng - 1);
    case RB_CT_F32: return vcycle_rb<float, float>(c, c.ng - 1);
    case RB_CT_F64: return vcycle_rb<double, double>(c, c.ng - 1);
    default: set_error("split V-cycle: coefficients not prepared"); return -1;
    }
}

}  // namespace dg
