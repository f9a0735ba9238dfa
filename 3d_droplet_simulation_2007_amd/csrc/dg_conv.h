// dg_conv.h -- BF16 3D convolution kernels (3x3x3, stride 1, zero padding)
// on channels-last tensors, MFMA implicit GEMM with FP32 accumulation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace dg {

typedef uint16_t bf16_t;  // storage of one bfloat16

// round-to-nearest-even FP32 -> BF16 (finite inputs)
__host__ __device__ inline bf16_t f2bf(float f)
{
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (bf16_t)(u >> 16);  // inf/nan pass through
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(bf16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

// A conv layer: cin -> cout, weights [coutp][27][cinp] with cinp a multiple
// of 8 and coutp a multiple of 16; taps ordered t = (dz+1)*9 + (dy+1)*3 + (dx+1).
struct ConvShape {
    int cin, cinp, cout, coutp;
};

// forward: y[v][co] = act(bias[co] + sum_{t,ci} x[v+off(t)][ci] w[co][t][ci]);
// x has xc channels per voxel (>= cinp, only the first cinp read), y is
// written with yc channels per voxel at channel offset yoff (bf16), or to yf
// (FP32, channel count yc) when yf != nullptr.  relu: 0 none, 1 ReLU.
int conv3d_fwd(int nx, int ny, int nz, const bf16_t *x, int xc, int cinp, const bf16_t *w, const float *bias,
               int cout, int coutp, int relu, bf16_t *y, int yc, int yoff, float *yf, hipStream_t s);

// weight gradient: gw[co][t][ci] = sum_v dz[v][co] x[v+off(t)][ci] (FP32,
// overwritten), bias gradient gb[co] = sum_v dz[v][co].  dz has coutp channels.
int conv3d_wgrad(int nx, int ny, int nz, const bf16_t *x, int xc, int cinp, const bf16_t *dz, int coutp, float *gw,
                 float *gb, float *scratch, size_t scratch_floats, hipStream_t s);

// Scratch floats conv3d_wgrad needs for this shape.
size_t conv3d_wgrad_scratch(int nx, int ny, int nz, int cinp, int coutp);

}  // namespace dg
