// dg_internal.h -- shared definitions of the MI355X droplet-step library.
//
// Layout in HBM: every field is one FP64 array of nx*ny*nz doubles with one
// ghost layer, x unit-stride, then y, then z (the reference's column-major
// (nx,ny,nz) order, main.f / input.h:33), so host<->device copies are flat.
// Indices in kernels are the reference's 1-based (i,j,k); the interior is
// 2..n-1.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace dg {

struct Grid {
    int nx, ny, nz;
    int sy, sz;  // strides of j and k
    __host__ __device__ inline int at(int i, int j, int k) const { return (i - 1) + sy * (j - 1) + sz * (k - 1); }
    __host__ __device__ inline size_t n() const { return (size_t)sz * nz; }
};

inline Grid make_grid(int nx, int ny, int nz)
{
    Grid g;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.sy = nx; g.sz = nx * ny;
    return g;
}

// error state of the library: the first failure is kept until read
void set_error(const char *fmt, ...);
const char *last_error();
int status();
// errors raised on the calling thread go to *sink (null: the library's
// error state); for worker threads, whose errors the caller reports later
void set_error_sink(std::string *sink);

#define DG_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            ::dg::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
            return -1;                                                                    \
        }                                                                                 \
    } while (0)

#define DG_CHECK_LAUNCH()                                                                 \
    do {                                                                                  \
        hipError_t e_ = hipGetLastError();                                                \
        if (e_ != hipSuccess) {                                                           \
            ::dg::set_error("%s:%d launch: %s", __FILE__, __LINE__, hipGetErrorString(e_)); \
            return -1;                                                                    \
        }                                                                                 \
    } while (0)

// interior launch shape: 64 lanes along unit-stride x, 4 rows of y, one z plane
constexpr int BX = 64, BY = 4;
inline dim3 interior_grid(const Grid &g)
{
    return dim3((unsigned)((g.nx - 2 + BX - 1) / BX), (unsigned)((g.ny - 2 + BY - 1) / BY), (unsigned)(g.nz - 2));
}
inline dim3 interior_block() { return dim3(BX, BY, 1); }

// The dispatcher deals a launch's workgroups to the 8 XCDs round-robin in
// launch order, and each XCD has its own L2.  xcd_tile() renumbers the
// launch so that XCD x runs one contiguous run of tiles (x fastest, then y,
// then z): the tiles beside a tile, whose halo rows and shared cache lines
// it reads, then sit in the same L2.  Returns the tile (bx, by, bz) of this
// workgroup, and its renumbered linear index through *lin.
__device__ inline uint3 xcd_tile(unsigned *lin = nullptr)
{
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned n = gx * gy * gridDim.z;
    unsigned id = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned per = n / 8;
    if (id < per * 8) id = (id % 8) * per + id / 8;
    if (lin) *lin = id;
    return make_uint3(id % gx, (id / gx) % gy, id / (gx * gy));
}

#define DG_INTERIOR_IJK(g)                                     \
    const uint3 tile_ = ::dg::xcd_tile();                       \
    const int i = (int)(tile_.x * blockDim.x + threadIdx.x) + 2; \
    const int j = (int)(tile_.y * blockDim.y + threadIdx.y) + 2; \
    const int k = (int)tile_.z + 2;                             \
    if (i > (g).nx - 1 || j > (g).ny - 1) return;

// 2-D launch over a face: (n1 x n2) points
inline dim3 face_grid(int n1, int n2) { return dim3((unsigned)((n1 + 255) / 256), (unsigned)n2, 1); }

// REAL*4 literals of the reference promoted to REAL*8
#define DG_F4(x) ((double)(float)(x))

}  // namespace dg
