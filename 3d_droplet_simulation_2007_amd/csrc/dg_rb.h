// dg_rb.h -- red-black split storage of the tolerance-mode V-cycle (FP32
// iterate in mode 2, FP64 iterate in mode 1).
//
// On every level above the coarsest, the unknowns u, the right-hand side f
// and the face coefficients a, b, c of the interior cells are stored per
// colour, colour = (i+j+k)&1 (1-based indices), in two halves of one array:
//   X[colour * half + ((k-2)*my + (j-2))*mh + ((i-2)>>1)],
//   mx = nx-2 (even), mh = mx/2, my = ny-2, mz = nz-2, half = mh*my*mz.
// Along a row (j,k) the cells of one colour are every other i, so a colour
// pass reads its own f and coefficients and the other colour's u and
// coefficients as contiguous streams and writes its own u contiguously: no
// half-used cache lines, and the other colour's unknowns are half the size
// of a full array.  Coefficients are stored in FP16 when their range allows
// (face values 1/rho of the two fluids) -- the V-cycle is a preconditioner
// of the FP64 PCG, which keeps the exact operator.
#pragma once

#include <hip/hip_fp16.h>

#include "dg_ctx.h"

namespace dg {

struct RbGeom {
    int nx, ny, nz;  // extents with ghosts
    int mh, my, mz;  // half row, interior y, interior z
    size_t half;     // cells of one colour
    __host__ __device__ inline size_t idx(int i, int j, int k) const
    {
        return ((size_t)(k - 2) * my + (j - 2)) * mh + ((i - 2) >> 1);
    }
    __host__ __device__ inline size_t at(int i, int j, int k) const
    {
        return (size_t)((i + j + k) & 1) * half + idx(i, j, k);
    }
    // the same offset in 32-bit arithmetic (levels with 2 * half < 2^31,
    // which rb_enabled requires of every split level)
    __device__ inline unsigned at32(int i, int j, int k) const
    {
        return (unsigned)((i + j + k) & 1) * (unsigned)half + (unsigned)(((k - 2) * my + (j - 2)) * mh + ((i - 2) >> 1));
    }
};

inline RbGeom rb_geom(const Grid &g)
{
    RbGeom r;
    r.nx = g.nx;
    r.ny = g.ny;
    r.nz = g.nz;
    r.mh = (g.nx - 2) / 2;
    r.my = g.ny - 2;
    r.mz = g.nz - 2;
    r.half = (size_t)r.mh * r.my * r.mz;
    return r;
}

// z accessors for the PCG kernels: the V-cycle result at interior (i,j,k)
template <typename T>
struct ZNatural {
    const T *z;
    Grid g;
    __device__ inline double operator()(int i, int j, int k, int o) const { return (double)z[o]; }
};
template <typename T>
struct ZSplit {
    const T *z;
    RbGeom r;
    __device__ inline double operator()(int i, int j, int k, int o) const { return (double)z[r.at32(i, j, k)]; }
};

// the split iterate / right-hand side arrays of a level for iterate type T
template <typename T>
inline T *rb_u_of(Level &L);
template <>
inline float *rb_u_of<float>(Level &L) { return L.rb_u; }
template <>
inline double *rb_u_of<double>(Level &L) { return L.rb_u64; }
template <typename T>
inline T *rb_f_of(Level &L);
template <>
inline float *rb_f_of<float>(Level &L) { return L.rb_f; }
template <>
inline double *rb_f_of<double>(Level &L) { return L.rb_f64; }

template <typename T>
inline T *rb_w_of(Level &L);
template <>
inline float *rb_w_of<float>(Level &L) { return L.rb_w; }
template <>
inline double *rb_w_of<double>(Level &L) { return L.rb_w64; }

// whether the split path serves this context's tolerance-mode solves
// (mode 2: FP32 iterate; mode 1: FP64 iterate with FP64 coefficients)
bool rb_enabled(const Ctx &c);
// the coefficient type (RB_CT_*) the split levels use for this context
int rb_coef_type(const Ctx &c);
// allocates the split arrays (and the coefficient planes of the type this
// solve uses) of every level above the coarsest
int rb_alloc(Ctx &c);
// rb_alloc, then the split coefficients of every level above the coarsest
// for this solve (after the coarse operators exist), except those written
// already (the finest by abcei, top_coef_ready; the coarse ones by
// acfromac, coarse_coef_ready); clears both flags
int rb_prepare(Ctx &c);
// finest-level split right-hand side (the iterate's type) from r (FP64,
// natural layout)
int rb_load_rhs(Ctx &c, const double *r);
// z = V-cycle(r) from a zero guess; r: FP64 finest-level residual (natural
// layout), or null when the split right-hand side already holds it; the
// result stays in the finest level's split u (rb_top_z)
int rb_vcycle(Ctx &c, const double *r);
template <typename T>
ZSplit<T> rb_top_z(Ctx &c)
{
    Level &F = c.lev[c.ng - 1];
    return ZSplit<T>{rb_u_of<T>(F), rb_geom(F.g)};
}

}  // namespace dg
