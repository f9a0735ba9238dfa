// dg_util.h -- launch helpers and reductions shared by the kernel files.
#pragma once

#include "dg_internal.h"

namespace dg {

// f(i,j,k) over the interior 2..n-1 of grid g
template <class F>
__global__ void __launch_bounds__(256) k_interior(Grid g, F f)
{
    DG_INTERIOR_IJK(g);
    f(i, j, k);
}

// the same in the dispatcher's own order (tiles dealt round-robin over the
// XCDs), for kernels whose work concentrates in a few tiles (the interface
// cells of PLIC and of the initial condition's sub-sampling): contiguous
// runs would leave most XCDs idle
template <class F>
__global__ void __launch_bounds__(256) k_interior_rr(Grid g, F f)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 2;
    const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) + 2;
    const int k = (int)blockIdx.z + 2;
    if (i > g.nx - 1 || j > g.ny - 1) return;
    f(i, j, k);
}

// f(a,b) over a 2-D index range [0,n1) x [0,n2)
template <class F>
__global__ void __launch_bounds__(256) k_face(int n1, int n2, F f)
{
    const int a = (int)(blockIdx.x * 256 + threadIdx.x);
    const int b = (int)blockIdx.y;
    if (a < n1 && b < n2) f(a, b);
}

// f(p) over [0,n)
template <class F>
__global__ void __launch_bounds__(256) k_linear(size_t n, F f)
{
    size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p < n) f(p);
}

template <class F>
inline int launch_interior(const Grid &g, hipStream_t s, F f)
{
    if (g.nx < 3 || g.ny < 3 || g.nz < 3) return 0;
    hipLaunchKernelGGL(k_interior<F>, interior_grid(g), interior_block(), 0, s, g, f);
    DG_CHECK_LAUNCH();
    return 0;
}

template <class F>
inline int launch_interior_rr(const Grid &g, hipStream_t s, F f)
{
    if (g.nx < 3 || g.ny < 3 || g.nz < 3) return 0;
    hipLaunchKernelGGL(k_interior_rr<F>, interior_grid(g), interior_block(), 0, s, g, f);
    DG_CHECK_LAUNCH();
    return 0;
}

template <class F>
inline int launch_face(int n1, int n2, hipStream_t s, F f)
{
    if (n1 <= 0 || n2 <= 0) return 0;
    hipLaunchKernelGGL(k_face<F>, face_grid(n1, n2), dim3(256), 0, s, n1, n2, f);
    DG_CHECK_LAUNCH();
    return 0;
}

template <class F>
inline int launch_linear(size_t n, hipStream_t s, F f)
{
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_linear<F>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, f);
    DG_CHECK_LAUNCH();
    return 0;
}

// max that keeps a NaN operand (fmax drops it), so a NaN anywhere in a
// maximum reduction (CFL, divergence) reaches the host
__host__ __device__ inline double max_nan(double a, double b) { return (a != a || a > b) ? a : b; }

// ---- block reductions (wave64 shuffles, then LDS across the 4 waves) ----
__device__ inline double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = max_nan(v, __shfl_xor(v, o, 64));
    return v;
}

template <bool IS_MAX>
__device__ inline double block_reduce256(double v, double *lds4)
{
    v = IS_MAX ? wave_max(v) : wave_sum(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) lds4[w] = v;
    __syncthreads();
    double r = lds4[0];
    for (int q = 1; q < 4; q++) r = IS_MAX ? max_nan(r, lds4[q]) : r + lds4[q];
    __syncthreads();
    return r;
}

// number of partial blocks used by the two-stage reductions
constexpr int RED_BLOCKS = 1024;

// stage 1: each of RED_BLOCKS blocks reduces a strided share of the
// interior; value(i,j,k) -> double.  Fixed order: deterministic run to run.
template <bool IS_MAX, class F>
__global__ void __launch_bounds__(256) k_reduce_interior(Grid g, F value, double *partial)
{
    __shared__ double lds4[4];
    const int mx = g.nx - 2, my = g.ny - 2, mz = g.nz - 2;
    const long long tot = (long long)mx * my * mz;
    double acc = IS_MAX ? -1.0 : 0.0;
    for (long long p = (long long)blockIdx.x * 256 + threadIdx.x; p < tot; p += (long long)gridDim.x * 256) {
        int i = (int)(p % mx) + 2;
        long long q = p / mx;
        int j = (int)(q % my) + 2;
        int k = (int)(q / my) + 2;
        double v = value(i, j, k);
        acc = IS_MAX ? max_nan(acc, v) : acc + v;
    }
    double r = block_reduce256<IS_MAX>(acc, lds4);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

template <bool IS_MAX>
__global__ void __launch_bounds__(256) k_reduce_final(const double *partial, int n, double *out)
{
    __shared__ double lds4[4];
    double acc = IS_MAX ? -1.0 : 0.0;
    for (int p = threadIdx.x; p < n; p += 256) acc = IS_MAX ? max_nan(acc, partial[p]) : acc + partial[p];
    double r = block_reduce256<IS_MAX>(acc, lds4);
    if (threadIdx.x == 0) *out = r;
}

template <bool IS_MAX, class F>
inline int reduce_interior(const Grid &g, hipStream_t s, double *partial, double *out, F value)
{
    const long long tot = (long long)(g.nx - 2) * (g.ny - 2) * (g.nz - 2);
    int nb = (int)((tot + 255) / 256);
    if (nb > RED_BLOCKS) nb = RED_BLOCKS;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL((k_reduce_interior<IS_MAX, F>), dim3(nb), dim3(256), 0, s, g, value, partial);
    DG_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_reduce_final<IS_MAX>, dim3(1), dim3(256), 0, s, partial, nb, out);
    DG_CHECK_LAUNCH();
    return 0;
}

}  // namespace dg
