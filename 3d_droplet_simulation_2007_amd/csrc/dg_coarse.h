// dg_coarse.h -- coarsest-level substitutions shared by the one-wave
// coarse-solve kernel and the split V-cycle's single-workgroup tail.
#pragma once

#include <hip/hip_runtime.h>

#include "dg_ctx.h"

namespace dg {

// u = -(L L^T)^-1 f on the coarsest level (natural layout), with M the
// Cholesky factor of -A from k_coarse_factor.  Called by every thread of a
// block (tid of nth); y: LDS staging of at least n doubles (n <= 512).
template <typename T>
__device__ inline void coarse_subst(const Grid &g, const double *M, const T *f, T *u, double *y, int tid, int nth)
{
    const int mx = g.nx - 2, my = g.ny - 2, mz = g.nz - 2, n = mx * my * mz;
    for (int row = tid; row < n; row += nth) {
        const int i = row % mx + 2, j = (row / mx) % my + 2, k = row / (mx * my) + 2;
        y[row] = -(double)f[g.at(i, j, k)];
    }
    __syncthreads();
    if (tid == 0) {
        for (int p = 0; p < n; p++) {
            double t = y[p];
            for (int q = 0; q < p; q++) t -= M[p * n + q] * y[q];
            y[p] = t / M[p * n + p];
        }
        for (int p = n - 1; p >= 0; p--) {
            double t = y[p];
            for (int q = p + 1; q < n; q++) t -= M[q * n + p] * y[q];
            y[p] = t / M[p * n + p];
        }
    }
    __syncthreads();
    for (int row = tid; row < n; row += nth) {
        const int i = row % mx + 2, j = (row / mx) % my + 2, k = row / (mx * my) + 2;
        u[g.at(i, j, k)] = (T)y[row];
    }
}

}  // namespace dg
