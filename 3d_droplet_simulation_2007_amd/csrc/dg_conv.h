// dg_conv.h -- BF16 3D convolution (3x3x3, stride 1, zero padding) and the
// other layers of the pressure-guess network, on channels-last tensors.
//
// Tensors: voxel-major, channel-minor bfloat16 ([v][c], v = (z*ny + y)*nx + x),
// no ghost layer; a tensor view is (pointer to its first channel, channel
// stride), so a layer can read or write a channel slice of a wider tensor
// (the skip concatenations).  Channel counts and slice offsets are multiples
// of 8 (16-byte vector loads).
//
// Convolution as an implicit GEMM on the matrix cores: M = voxels, N = output
// channels, K = 27 taps x input channels, v_mfma_f32_16x16x32_bf16 tiles with
// FP32 accumulation; weight-gradient GEMM with K = voxels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace dg {

typedef uint16_t bf16_t;  // storage of one bfloat16

// round-to-nearest-even FP32 -> BF16 (NaN stays NaN)
__host__ __device__ inline bf16_t f2bf(float f)
{
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (bf16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
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

// Master weights of one layer: FP32 [coutp][27][cinp], taps ordered
// t = (dz+1)*9 + (dy+1)*3 + (dx+1); padded rows/columns are zero.
// Forward B-operand packing: wB[g][coutp][8] with g = t*(cinp/8) + c8,
// wB[g][co][e] = w[co][t][c8*8+e].
int conv_pack_fwd(const float *w, int cinp, int coutp, bf16_t *wB, hipStream_t s);
// Data-gradient packing (the forward convolution of dy with the flipped,
// transposed kernel): wD[g][cinp][8] with g = t*(coutp/8) + c8,
// wD[g][ci][e] = w[c8*8+e][26-t][ci].
int conv_pack_dgrad(const float *w, int cinp, int coutp, bf16_t *wD, hipStream_t s);

// y[v][co] = act(bias[co] + sum_{t,ci} x[v+off(t)][ci] * w[co][t][ci]) for
// co < cout (nothing is written for cout <= co < coutp).  x: cinp channels,
// stride xs.  Output bf16 (y, stride ys) or, when yf != nullptr, FP32 (yf,
// stride yfs).  relu: 1 ReLU, 0 identity.  bias may be null.  coutp is a
// multiple of 16 up to 128; cinp a multiple of 8.
int conv3d_fwd(int nx, int ny, int nz, const bf16_t *x, int xs, int cinp, const bf16_t *wB, int coutp,
               const float *bias, int cout, int relu, bf16_t *y, int ys, float *yf, int yfs, hipStream_t s);

// dx[v][ci] = sum_{t,co} dy[v-off(t)][co] * w[co][t][ci], ci < cinp:
// conv3d_fwd of dy (coutp channels) with conv_pack_dgrad's weights.
inline int conv3d_dgrad(int nx, int ny, int nz, const bf16_t *dy, int dys, int coutp, const bf16_t *wD, int cinp,
                        bf16_t *dx, int dxs, hipStream_t s)
{
    return conv3d_fwd(nx, ny, nz, dy, dys, coutp, wD, cinp, nullptr, cinp, 0, dx, dxs, nullptr, 0, s);
}

// Scratch floats conv3d_wgrad needs for this shape.
size_t conv3d_wgrad_scratch(int nx, int ny, int nz, int cinp, int coutp);
// gw[co][t][ci] = sum_v dy[v][co] * x[v+off(t)][ci] (FP32, overwritten, all
// coutp x 27 x cinp entries) and, when gb != nullptr, gb[co] = sum_v dy[v][co].
// Deterministic: fixed partition of the voxels, partial sums reduced in order.
int conv3d_wgrad(int nx, int ny, int nz, const bf16_t *x, int xs, int cinp, const bf16_t *dy, int dys, int coutp,
                 float *gw, float *gb, float *scratch, size_t scratch_floats, hipStream_t s);

// 2x2x2 average pooling: out (nx/2, ny/2, nz/2) from in (nx, ny, nz), c channels.
int pool2_fwd(int nx, int ny, int nz, int c, const bf16_t *in, int is, bf16_t *out, int os, hipStream_t s);
// its adjoint, accumulated: din[child] += dout[parent] / 8.
int pool2_bwd_add(int nx, int ny, int nz, int c, const bf16_t *dout, int ds, bf16_t *din, int is, hipStream_t s);
// nearest 2x upsampling: out (nx, ny, nz) from in (nx/2, ny/2, nz/2).
int up2_fwd(int nx, int ny, int nz, int c, const bf16_t *in, int is, bf16_t *out, int os, hipStream_t s);
// its adjoint: din[parent] = sum of dout over the 8 children.
int up2_bwd(int nx, int ny, int nz, int c, const bf16_t *dout, int ds, bf16_t *din, int is, hipStream_t s);
// g[v][c] = y[v][c] > 0 ? g[v][c] : 0 (ReLU adjoint from the stored output).
int relu_bwd(size_t nvox, int c, const bf16_t *y, int ys, bf16_t *g, int gs, hipStream_t s);

// Adam (FP32): m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;
// p -= lr_t * m / (sqrt(v) + eps_t) with the bias corrections folded into
// lr_t and eps_t by the caller.
int adam_update(size_t n, float *p, const float *g, float *m, float *v, float b1, float b2, float lr_t, float eps_t,
                hipStream_t s);

}  // namespace dg
