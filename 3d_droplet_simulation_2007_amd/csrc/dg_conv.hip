// dg_conv.hip -- BF16 MFMA 3D convolution kernels and the elementwise /
// resampling layers of the pressure-guess network (gfx950).

#include <hip/hip_runtime.h>

#include "dg_conv.h"
#include "dg_internal.h"
#include "dg_prof.h"

namespace dg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline bf16x8 ld8(const bf16_t *p)
{
    const uint4 u = *reinterpret_cast<const uint4 *>(p);
    return __builtin_bit_cast(bf16x8, u);
}
__device__ inline bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0u, 0u, 0u, 0u)); }

// Blocks are dealt round-robin to the 8 XCDs; give each XCD a contiguous
// range of voxel tiles so that the halo rows shared by neighbouring tiles
// hit the same L2.  Bijective when nb is a multiple of 8 (grids are rounded
// up to one), identity otherwise.
__device__ inline int xcd_tile(int b, int nb) { return (nb & 7) ? b : (b & 7) * (nb >> 3) + (b >> 3); }

// ---- forward / data-gradient convolution ------------------------------------
//
// One wave computes 32 voxels (two 16-row M tiles) x NT*16 output channels.
// K loop: groups of 8 input channels of one tap, four groups (K = 32) per
// MFMA; lane (row = l%16, kg = l/16) loads 8 channels of voxel `row`'s
// neighbour for group 4*step+kg (A) and 8 weights of output channel `row`
// for the same group (B).  The K order inside an MFMA is the same for A and
// B, so any lane-to-k assignment gives the same dot product; the output
// tile is D[voxel 4*kg+r][channel row] in lane l's register r.
template <int NT>
__global__ void __launch_bounds__(256) k_conv_fwd(int nx, int ny, int nz, const bf16_t *__restrict__ x, int xs, int cg,
                                                  const bf16_t *__restrict__ wB, int coutp,
                                                  const float *__restrict__ bias, int cout, int relu,
                                                  bf16_t *__restrict__ y, int ys, float *__restrict__ yf, int yfs)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, kg = lane >> 4;
    const long long V = (long long)nx * ny * nz;
    const int blk = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const long long v0 = ((long long)blk * 4 + wave) * 32;
    if (v0 >= V) return;
    int px[2], py[2], pz[2];
    bool pin[2];
#pragma unroll
    for (int m = 0; m < 2; m++) {
        long long v = v0 + m * 16 + row;
        pin[m] = v < V;
        if (!pin[m]) v = 0;
        px[m] = (int)(v % nx);
        const long long q = v / nx;
        py[m] = (int)(q % ny);
        pz[m] = (int)(q / ny);
    }
    f32x4 acc[2][NT];
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int n = 0; n < NT; n++) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int G = 27 * cg;
    int g = kg, t = kg / cg, c8 = kg - (kg / cg) * cg;
    for (int g0 = 0; g0 < G; g0 += 4, g += 4) {
        const bool gv = g < G;
        const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
        bf16x8 a[2], b[NT];
#pragma unroll
        for (int m = 0; m < 2; m++) {
            const int xx = px[m] + dx, yy = py[m] + dy, zz = pz[m] + dz;
            const bool ok = gv && pin[m] && (unsigned)xx < (unsigned)nx && (unsigned)yy < (unsigned)ny &&
                            (unsigned)zz < (unsigned)nz;
            a[m] = ok ? ld8(x + ((size_t)((long long)zz * ny + yy) * nx + xx) * xs + c8 * 8) : zero8();
        }
        const bf16_t *wg = wB + ((size_t)(gv ? g : 0) * coutp + row) * 8;
#pragma unroll
        for (int n = 0; n < NT; n++) b[n] = gv ? ld8(wg + n * 128) : zero8();
#pragma unroll
        for (int m = 0; m < 2; m++)
#pragma unroll
            for (int n = 0; n < NT; n++)
                acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b[n], acc[m][n], 0, 0, 0);
        c8 += 4;
        while (c8 >= cg) {
            c8 -= cg;
            t++;
        }
    }
#pragma unroll
    for (int n = 0; n < NT; n++) {
        const int co = n * 16 + row;
        if (co >= cout) continue;
        const float bv = bias ? bias[co] : 0.f;
#pragma unroll
        for (int m = 0; m < 2; m++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const long long v = v0 + m * 16 + kg * 4 + r;
                if (v >= V) continue;
                float val = acc[m][n][r] + bv;
                if (relu) val = fmaxf(val, 0.f);
                if (yf)
                    yf[(size_t)v * yfs + co] = val;
                else
                    y[(size_t)v * ys + co] = f2bf(val);
            }
    }
}

// The same convolution with the input staged in LDS: a workgroup owns a
// tile of TX x 2R x 2 voxels (TX = 32, or 16 on grids of 16 to 31 cells in
// x; each wave R rows of TX voxels: TX / 16 * R M tiles of 16 voxels; R = 2
// up to 48 input channels, else 1) and first copies the tile's input with
// its one-voxel halo, (TX+2) x (2R+2) x 4 voxels of all cinp channels, from
// HBM into LDS with 16-byte loads (zeros outside the grid).
// The K loop is unrolled at compile time (CG = cinp / 8 channel groups, NT =
// coutp / 16 output tiles): each step reads the lane's A offset from a small
// LDS table (k group -> tap and channel group), its A fragments from the
// halo tile at constant distances, and its B fragment (packed weights, from
// global memory: small, cached) D steps ahead of its use, once for all of
// the wave's M tiles.  Same K order, same MFMAs: the same bits.
constexpr int CF_TZ = 2;

template <int R, int TX>
struct CfTile {
    static constexpr int MPR = TX / 16, TY = 2 * R, HX = TX + 2, HY = TY + 2, HZ = CF_TZ + 2, HN = HX * HY * HZ;
};

template <int CG, int TX>
struct CfShape {
    static constexpr int R = CG <= 6 ? 2 : 1, G = 27 * CG, NS = (G + 3) / 4, NSP = (NS + 3) / 4 * 4,
                         REM = G - 4 * (NS - 1);
    using TL = CfTile<R, TX>;
    static constexpr size_t lds = (size_t)TL::HN * CG * 16 + (size_t)4 * NSP * 4;
};

// halo-tile offset, in voxels, of tap t (x fastest, then y, then z)
template <int HX, int HY>
__device__ constexpr int cf_toff(int t)
{
    return (t / 9 - 1) * HX * HY + ((t / 3) % 3 - 1) * HX + (t % 3 - 1);
}

template <int NT, int CG, int TX>
__global__ void __launch_bounds__(256) k_conv_fwd_lds(int nx, int ny, int nz, const bf16_t *__restrict__ x, int xs,
                                                      const bf16_t *__restrict__ wB, const float *__restrict__ bias,
                                                      int cout, int relu, bf16_t *__restrict__ y, int ys,
                                                      float *__restrict__ yf, int yfs, int ntx, int nty)
{
    using S = CfShape<CG, TX>;
    using TL = typename S::TL;
    constexpr int R = S::R, MPR = TL::MPR, MT = MPR * R, NS = S::NS, NSP = S::NSP, REM = S::REM;
    constexpr int COUTP = 16 * NT;
    constexpr int D = NT >= 4 ? 2 : 4;  // B prefetch distance in steps
    extern __shared__ __align__(16) unsigned char cf_lds[];
    uint4 *const hal = reinterpret_cast<uint4 *>(cf_lds);  // [voxel of the halo tile][channel group]
    int *const tab = reinterpret_cast<int *>(hal + TL::HN * CG);  // [k group mod 4][step]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, kg = lane >> 4;
    const int blk = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int tx = blk % ntx, ty = (blk / ntx) % nty, tz = blk / (ntx * nty);
    const int x0 = tx * TX, y0 = ty * TL::TY, z0 = tz * CF_TZ;
    if (z0 >= nz) return;  // the grid's rounding to a multiple of 8 blocks
    // offset table: k group g = 4 s + q reads channel group g % CG of tap
    // g / CG, at this many uint4 from the output voxel's own first group
    for (int e = (int)threadIdx.x; e < 4 * NSP; e += 256) {
        const int q = e / NSP, s = e % NSP, g = min(4 * s + q, S::G - 1);
        const int t = g / CG;
        tab[e] = cf_toff<TL::HX, TL::HY>(t) * CG + (g - t * CG);
    }
    // the halo tile, eight 16-byte loads in flight per thread before their
    // stores; a place outside the grid loads voxel 0 and stores zeros
    constexpr int UB = 8, NE = TL::HN * CG;
    for (int e0 = (int)threadIdx.x; e0 < NE; e0 += 256 * UB) {
        uint4 v[UB];
        bool in[UB];
#pragma unroll
        for (int u = 0; u < UB; u++) {
            const int e = min(e0 + u * 256, NE - 1);
            const int hv = e / CG, c8 = e - hv * CG;
            const int hx = hv % TL::HX, hy = (hv / TL::HX) % TL::HY, hz = hv / (TL::HX * TL::HY);
            const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
            in[u] = (unsigned)gx < (unsigned)nx && (unsigned)gy < (unsigned)ny && (unsigned)gz < (unsigned)nz;
            const size_t src = in[u] ? ((size_t)((long long)gz * ny + gy) * nx + gx) * xs + c8 * 8 : 0;
            v[u] = *reinterpret_cast<const uint4 *>(x + src);
        }
#pragma unroll
        for (int u = 0; u < UB; u++)
            if (e0 + u * 256 < NE) hal[e0 + u * 256] = in[u] ? v[u] : make_uint4(0u, 0u, 0u, 0u);
    }
    // B fragments of the first D steps while the tile lands
    const bf16_t *const wl = wB + ((size_t)kg * COUTP + row) * 8;  // step s adds s * 4 * COUTP * 8
    auto load_b = [&](int s, bf16x8 *b) {
        const bool ok = s < NS - 1 || REM == 4 || kg < REM;
#pragma unroll
        for (int n = 0; n < NT; n++) b[n] = ok ? ld8(wl + (size_t)(ok ? s : 0) * 4 * COUTP * 8 + n * 128) : zero8();
    };
    bf16x8 bq[D][NT];
#pragma unroll
    for (int s = 0; s < D; s++)
        if (s < NS) load_b(s, bq[s]);
    __syncthreads();
    // the wave's rows are tile rows wave * R + q, q < R, which share z (TY
    // is even): row q starts q * HX halo voxels after row 0
    const int tr0 = wave * R, ry = tr0 % TL::TY, rz = tr0 / TL::TY;
    const uint4 *const ab = hal + (((rz + 1) * TL::HY + (ry + 1)) * TL::HX + 1 + row) * CG;
    const int4 *const tp = reinterpret_cast<const int4 *>(tab + kg * NSP);
    f32x4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int n = 0; n < NT; n++) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
    int4 tv = tp[0];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int off = (s & 3) == 0 ? tv.x : (s & 3) == 1 ? tv.y : (s & 3) == 2 ? tv.z : tv.w;
        if ((s & 3) == 3 && s + 1 < NS) tv = tp[(s + 1) >> 2];
        const bool ok = s < NS - 1 || REM == 4 || kg < REM;
        bf16x8 a[MT];
#pragma unroll
        for (int m = 0; m < MT; m++) {
            const bf16x8 av = __builtin_bit_cast(bf16x8, ab[off + ((m / MPR) * TL::HX + (m % MPR) * 16) * CG]);
            a[m] = ok ? av : zero8();
        }
        bf16x8 *const b = bq[s % D];
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int n = 0; n < NT; n++) acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b[n], acc[m][n], 0, 0, 0);
        if (s + D < NS) load_b(s + D, b);
    }
#pragma unroll
    for (int n = 0; n < NT; n++) {
        const int co = n * 16 + row;
        if (co >= cout) continue;
        const float bv = bias ? bias[co] : 0.f;
#pragma unroll
        for (int m = 0; m < MT; m++) {
            const int tr = tr0 + m / MPR, gyw = y0 + tr % TL::TY, gzw = z0 + tr / TL::TY;
            if (gyw >= ny || gzw >= nz) continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int gx = x0 + (m % MPR) * 16 + kg * 4 + r;
                if (gx >= nx) continue;
                const size_t v = ((size_t)gzw * ny + gyw) * nx + gx;
                float val = acc[m][n][r] + bv;
                if (relu) val = fmaxf(val, 0.f);
                if (yf)
                    yf[v * yfs + co] = val;
                else
                    y[v * ys + co] = f2bf(val);
            }
        }
    }
}

template <int NT, int CG, int TX>
static int launch_conv_fwd_lds(int nx, int ny, int nz, const bf16_t *x, int xs, const bf16_t *wB, const float *bias,
                               int cout, int relu, bf16_t *y, int ys, float *yf, int yfs, hipStream_t s)
{
    using S = CfShape<CG, TX>;
    using TL = typename S::TL;
    const int ntx = (nx + TX - 1) / TX, nty = (ny + TL::TY - 1) / TL::TY, ntz = (nz + CF_TZ - 1) / CF_TZ;
    long long nbl = (long long)ntx * nty * ntz;
    nbl = (nbl + 7) / 8 * 8;
    if (S::lds > 48 * 1024)
        DG_HIP(hipFuncSetAttribute((const void *)k_conv_fwd_lds<NT, CG, TX>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)S::lds));
    hipLaunchKernelGGL((k_conv_fwd_lds<NT, CG, TX>), dim3((unsigned)nbl), dim3(256), S::lds, s, nx, ny, nz, x, xs, wB,
                       bias, cout, relu, y, ys, yf, yfs, ntx, nty);
    return 0;
}

template <int CG, int TX>
static int conv_fwd_lds_nt(int nx, int ny, int nz, const bf16_t *x, int xs, const bf16_t *wB, int coutp,
                           const float *bias, int cout, int relu, bf16_t *y, int ys, float *yf, int yfs, hipStream_t s)
{
    switch (coutp) {
    case 16: return launch_conv_fwd_lds<1, CG, TX>(nx, ny, nz, x, xs, wB, bias, cout, relu, y, ys, yf, yfs, s);
    case 32: return launch_conv_fwd_lds<2, CG, TX>(nx, ny, nz, x, xs, wB, bias, cout, relu, y, ys, yf, yfs, s);
    default: return launch_conv_fwd_lds<4, CG, TX>(nx, ny, nz, x, xs, wB, bias, cout, relu, y, ys, yf, yfs, s);
    }
}

// 16-voxel rows also on wider grids where they measured faster (4, 8, 12 and
// 16 channel groups: the smaller halo tile fits more workgroups per CU; with
// 1, 2 or 6 groups the 32-voxel rows' B reuse over more M tiles wins)
template <int CG>
static int conv_fwd_lds_tx(int nx, int ny, int nz, const bf16_t *x, int xs, const bf16_t *wB, int coutp,
                           const float *bias, int cout, int relu, bf16_t *y, int ys, float *yf, int yfs, hipStream_t s)
{
    if constexpr (CG <= 2 || CG == 6)
        if (nx >= 32) return conv_fwd_lds_nt<CG, 32>(nx, ny, nz, x, xs, wB, coutp, bias, cout, relu, y, ys, yf, yfs, s);
    return conv_fwd_lds_nt<CG, 16>(nx, ny, nz, x, xs, wB, coutp, bias, cout, relu, y, ys, yf, yfs, s);
}

int conv3d_fwd(int nx, int ny, int nz, const bf16_t *x, int xs, int cinp, const bf16_t *wB, int coutp,
               const float *bias, int cout, int relu, bf16_t *y, int ys, float *yf, int yfs, hipStream_t s)
{
    if (cinp <= 0 || (cinp & 7) || coutp <= 0 || (coutp & 15) || coutp > 128 || cout > coutp || xs < cinp ||
        (xs & 7) || (!yf && (ys < cout || (ys & 7))) || (yf && yfs < cout)) {
        set_error("conv3d_fwd: unsupported shape cinp=%d xs=%d coutp=%d cout=%d ys=%d", cinp, xs, coutp, cout, ys);
        return -1;
    }
    if (((uintptr_t)x & 15) || ((uintptr_t)wB & 15)) {
        set_error("conv3d_fwd: operands must be 16-byte aligned");
        return -1;
    }
    const long long V = (long long)nx * ny * nz;
    if (V <= 0) return 0;
    const int cg = cinp / 8;
    // the LDS-staged form where a row of 16 or 32 voxels is mostly inside
    // the grid and the halo tile fits the LDS (up to 128 input channels:
    // 139 KB)
    if (nx >= 16 && (cg <= 2 || cg == 4 || cg == 6 || cg == 8 || cg == 12 || cg == 16) &&
        (coutp == 16 || coutp == 32 || coutp == 64)) {
        prof_begin(PK_CONV_FWD, s);
        int rc;
        switch (cg) {
        case 1: rc = conv_fwd_lds_tx<1>(nx, ny, nz, x, xs, wB, coutp, bias, cout, relu, y, ys, yf, yfs, s); break;
        case 2: rc = conv_fwd_lds_tx<2>(nx, ny, nz, x, xs, wB, coutp, bias, cout, relu, y, ys, yf, yfs, s); break;
        case 4: rc = conv_fwd_lds_tx<4>(nx, ny, nz, x, xs, wB, coutp, bias, cout, relu, y, ys, yf, yfs, s); break;
        case 6: rc = conv_fwd_lds_tx<6>(nx, ny, nz, x, xs, wB, coutp, bias, cout, relu, y, ys, yf, yfs, s); break;
        case 8: rc = conv_fwd_lds_tx<8>(nx, ny, nz, x, xs, wB, coutp, bias, cout, relu, y, ys, yf, yfs, s); break;
        case 12: rc = conv_fwd_lds_tx<12>(nx, ny, nz, x, xs, wB, coutp, bias, cout, relu, y, ys, yf, yfs, s); break;
        default: rc = conv_fwd_lds_tx<16>(nx, ny, nz, x, xs, wB, coutp, bias, cout, relu, y, ys, yf, yfs, s); break;
        }
        if (rc) return -1;
        DG_CHECK_LAUNCH();
        prof_end(PK_CONV_FWD, s, 2.0 * (double)V * 27.0 * (double)cinp * (double)cout);
        return 0;
    }
    long long nb = (V + 127) / 128;
    nb = (nb + 7) / 8 * 8;
    const dim3 grid((unsigned)nb), block(256);
    prof_begin(PK_CONV_FWD, s);
    switch (coutp / 16) {
    case 1: hipLaunchKernelGGL(k_conv_fwd<1>, grid, block, 0, s, nx, ny, nz, x, xs, cg, wB, coutp, bias, cout, relu, y, ys, yf, yfs); break;
    case 2: hipLaunchKernelGGL(k_conv_fwd<2>, grid, block, 0, s, nx, ny, nz, x, xs, cg, wB, coutp, bias, cout, relu, y, ys, yf, yfs); break;
    case 3: hipLaunchKernelGGL(k_conv_fwd<3>, grid, block, 0, s, nx, ny, nz, x, xs, cg, wB, coutp, bias, cout, relu, y, ys, yf, yfs); break;
    case 4: hipLaunchKernelGGL(k_conv_fwd<4>, grid, block, 0, s, nx, ny, nz, x, xs, cg, wB, coutp, bias, cout, relu, y, ys, yf, yfs); break;
    case 6: hipLaunchKernelGGL(k_conv_fwd<6>, grid, block, 0, s, nx, ny, nz, x, xs, cg, wB, coutp, bias, cout, relu, y, ys, yf, yfs); break;
    case 8: hipLaunchKernelGGL(k_conv_fwd<8>, grid, block, 0, s, nx, ny, nz, x, xs, cg, wB, coutp, bias, cout, relu, y, ys, yf, yfs); break;
    default:
        set_error("conv3d_fwd: coutp=%d not instantiated (16, 32, 48, 64, 96, 128)", coutp);
        return -1;
    }
    DG_CHECK_LAUNCH();
    // algorithmic FLOPs: a multiply and an add per tap, input channel (as
    // padded) and output channel of every voxel
    prof_end(PK_CONV_FWD, s, 2.0 * (double)V * 27.0 * (double)cinp * (double)cout);
    return 0;
}

// ---- weight packing -------------------------------------------------------------

__global__ void k_pack_fwd(const float *w, int cinp, int coutp, bf16_t *wB)
{
    const int cg = cinp / 8;
    const size_t n = (size_t)27 * cg * coutp * 8;
    for (size_t p = (size_t)blockIdx.x * 256 + threadIdx.x; p < n; p += (size_t)gridDim.x * 256) {
        const int e = (int)(p & 7);
        const size_t q = p >> 3;
        const int co = (int)(q % coutp);
        const int g = (int)(q / coutp);
        const int t = g / cg, c8 = g % cg;
        wB[p] = f2bf(w[((size_t)co * 27 + t) * cinp + c8 * 8 + e]);
    }
}

__global__ void k_pack_dgrad(const float *w, int cinp, int coutp, bf16_t *wD)
{
    const int cgo = coutp / 8;
    const size_t n = (size_t)27 * cgo * cinp * 8;
    for (size_t p = (size_t)blockIdx.x * 256 + threadIdx.x; p < n; p += (size_t)gridDim.x * 256) {
        const int e = (int)(p & 7);
        const size_t q = p >> 3;
        const int ci = (int)(q % cinp);
        const int g = (int)(q / cinp);
        const int t = g / cgo, c8 = g % cgo;
        wD[p] = f2bf(w[((size_t)(c8 * 8 + e) * 27 + (26 - t)) * cinp + ci]);
    }
}

static unsigned grid_for(size_t n)
{
    size_t b = (n + 255) / 256;
    if (b > 65535) b = 65535;
    if (b < 1) b = 1;
    return (unsigned)b;
}

int conv_pack_fwd(const float *w, int cinp, int coutp, bf16_t *wB, hipStream_t s)
{
    hipLaunchKernelGGL(k_pack_fwd, dim3(grid_for((size_t)27 * cinp * coutp)), dim3(256), 0, s, w, cinp, coutp, wB);
    DG_CHECK_LAUNCH();
    return 0;
}

int conv_pack_dgrad(const float *w, int cinp, int coutp, bf16_t *wD, hipStream_t s)
{
    hipLaunchKernelGGL(k_pack_dgrad, dim3(grid_for((size_t)27 * cinp * coutp)), dim3(256), 0, s, w, cinp, coutp, wD);
    DG_CHECK_LAUNCH();
    return 0;
}

// ---- weight gradient ----------------------------------------------------------------
//
// GEMM gw[co][col] = sum_v dy[v][co] * X[v][col], col = t*cinp + ci, with
// K = voxels.  A wave owns MT row tiles (all of coutp) x NJ column tiles
// and a contiguous voxel chunk; per 32 voxels lane (row, kg) gathers the
// 8 voxels kg*8..kg*8+7 of dy (channel mt*16+row) and of x's tap-shifted
// neighbour (column tile*16+row).  Partial tiles go to part[chunk][co][col]
// and are summed over the chunks in a fixed order.

constexpr int WG_NJ = 4;

template <int MT>
__global__ void __launch_bounds__(256) k_conv_wgrad(int nx, int ny, int nz, const bf16_t *__restrict__ x, int xs,
                                                    int cinp, const bf16_t *__restrict__ dy, int dys, int ncoltiles,
                                                    int ngroups, int nchunks, long long chunk_vox,
                                                    float *__restrict__ part)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, kg = lane >> 4;
    const long long wid = (long long)blockIdx.x * 4 + wave;
    if (wid >= (long long)ngroups * nchunks) return;
    const int grp = (int)(wid % ngroups);
    const int chunk = (int)(wid / ngroups);
    const long long V = (long long)nx * ny * nz;
    const long long vbeg = (long long)chunk * chunk_vox;
    long long vend = vbeg + chunk_vox;
    if (vend > V) vend = V;
    const int ncols = 27 * cinp;
    int cdx[WG_NJ], cdy[WG_NJ], cdz[WG_NJ], cci[WG_NJ];
    bool cok[WG_NJ];
#pragma unroll
    for (int j = 0; j < WG_NJ; j++) {
        const int tile = grp * WG_NJ + j;
        const int col = tile * 16 + row;
        cok[j] = tile < ncoltiles && col < ncols;
        const int t = cok[j] ? col / cinp : 0;
        cci[j] = cok[j] ? col % cinp : 0;
        cdz[j] = t / 9 - 1;
        cdy[j] = (t / 3) % 3 - 1;
        cdx[j] = t % 3 - 1;
    }
    f32x4 acc[MT][WG_NJ];
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int j = 0; j < WG_NJ; j++) acc[m][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (long long vb = vbeg; vb < vend; vb += 32) {
        // coordinates of this lane's first voxel; the other 7 follow along x
        long long v = vb + kg * 8;
        int vx = (int)(v % nx);
        long long q = v / nx;
        int vy = (int)(q % ny), vz = (int)(q / ny);
        uint16_t av[MT][8], bv[WG_NJ][8];
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const bool vin = v + e < vend;
#pragma unroll
            for (int m = 0; m < MT; m++) av[m][e] = vin ? dy[(size_t)(v + e) * dys + m * 16 + row] : (uint16_t)0;
#pragma unroll
            for (int j = 0; j < WG_NJ; j++) {
                const int xx = vx + cdx[j], yy = vy + cdy[j], zz = vz + cdz[j];
                const bool ok = vin && cok[j] && (unsigned)xx < (unsigned)nx && (unsigned)yy < (unsigned)ny &&
                                (unsigned)zz < (unsigned)nz;
                bv[j][e] = ok ? x[((size_t)((long long)zz * ny + yy) * nx + xx) * xs + cci[j]] : (uint16_t)0;
            }
            if (++vx == nx) {
                vx = 0;
                if (++vy == ny) {
                    vy = 0;
                    vz++;
                }
            }
        }
        bf16x8 a[MT], b[WG_NJ];
#pragma unroll
        for (int m = 0; m < MT; m++) {
            uint4 u;
            u.x = av[m][0] | ((uint32_t)av[m][1] << 16);
            u.y = av[m][2] | ((uint32_t)av[m][3] << 16);
            u.z = av[m][4] | ((uint32_t)av[m][5] << 16);
            u.w = av[m][6] | ((uint32_t)av[m][7] << 16);
            a[m] = __builtin_bit_cast(bf16x8, u);
        }
#pragma unroll
        for (int j = 0; j < WG_NJ; j++) {
            uint4 u;
            u.x = bv[j][0] | ((uint32_t)bv[j][1] << 16);
            u.y = bv[j][2] | ((uint32_t)bv[j][3] << 16);
            u.z = bv[j][4] | ((uint32_t)bv[j][5] << 16);
            u.w = bv[j][6] | ((uint32_t)bv[j][7] << 16);
            b[j] = __builtin_bit_cast(bf16x8, u);
        }
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int j = 0; j < WG_NJ; j++)
                acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b[j], acc[m][j], 0, 0, 0);
    }
    const int coutp = MT * 16;
    float *pc = part + (size_t)chunk * coutp * ncols;
#pragma unroll
    for (int j = 0; j < WG_NJ; j++) {
        if (!cok[j]) continue;
        const int col = (grp * WG_NJ + j) * 16 + row;
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int r = 0; r < 4; r++) pc[(size_t)(m * 16 + kg * 4 + r) * ncols + col] = acc[m][j][r];
    }
}

__global__ void k_sum_chunks(const float *part, size_t n, int nchunks, float *out)
{
    for (size_t p = (size_t)blockIdx.x * 256 + threadIdx.x; p < n; p += (size_t)gridDim.x * 256) {
        float acc = 0.f;
        for (int c = 0; c < nchunks; c++) acc += part[(size_t)c * n + p];
        out[p] = acc;
    }
}

// bias gradient: block b sums a voxel chunk for every channel, then the
// chunks are summed in order
__global__ void __launch_bounds__(256) k_bias_part(long long V, const bf16_t *dy, int dys, int coutp, long long chunk_vox,
                                                   float *part)
{
    __shared__ float lds[256];
    const int lanes = 256 / coutp;  // voxel lanes per channel
    const int co = threadIdx.x % coutp, vl = threadIdx.x / coutp;
    const long long vbeg = (long long)blockIdx.x * chunk_vox;
    long long vend = vbeg + chunk_vox;
    if (vend > V) vend = V;
    float acc = 0.f;
    if (vl < lanes)
        for (long long v = vbeg + vl; v < vend; v += lanes) acc += bf2f(dy[(size_t)v * dys + co]);
    lds[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < coutp) {
        float s = 0.f;
        for (int q = 0; q < lanes; q++) s += lds[q * coutp + threadIdx.x];
        part[(size_t)blockIdx.x * coutp + threadIdx.x] = s;
    }
}

static void wgrad_plan(long long V, int cinp, int coutp, int &ngroups, int &nchunks, long long &chunk_vox)
{
    const int ncoltiles = (27 * cinp + 15) / 16;
    ngroups = (ncoltiles + WG_NJ - 1) / WG_NJ;
    // about 8192 waves, at least 256 voxels per wave
    long long want = 8192 / ngroups;
    if (want < 1) want = 1;
    long long maxc = (V + 255) / 256;
    if (want > maxc) want = maxc;
    if (want < 1) want = 1;
    chunk_vox = (V + want - 1) / want;
    chunk_vox = (chunk_vox + 31) / 32 * 32;
    nchunks = (int)((V + chunk_vox - 1) / chunk_vox);
    if (nchunks < 1) nchunks = 1;
}

size_t conv3d_wgrad_scratch(int nx, int ny, int nz, int cinp, int coutp)
{
    int ngroups, nchunks;
    long long chunk_vox;
    const long long V = (long long)nx * ny * nz;
    wgrad_plan(V, cinp, coutp, ngroups, nchunks, chunk_vox);
    const size_t wpart = (size_t)nchunks * coutp * 27 * cinp;
    const long long bchunks = (V + 4095) / 4096;
    return wpart + (size_t)bchunks * coutp + 64;
}

int conv3d_wgrad(int nx, int ny, int nz, const bf16_t *x, int xs, int cinp, const bf16_t *dy, int dys, int coutp,
                 float *gw, float *gb, float *scratch, size_t scratch_floats, hipStream_t s)
{
    if (cinp <= 0 || (cinp & 7) || coutp <= 0 || (coutp & 15) || coutp > 64 || xs < cinp || dys < coutp) {
        set_error("conv3d_wgrad: unsupported shape cinp=%d coutp=%d", cinp, coutp);
        return -1;
    }
    const long long V = (long long)nx * ny * nz;
    if (V <= 0) {
        set_error("conv3d_wgrad: empty grid");
        return -1;
    }
    if (scratch_floats < conv3d_wgrad_scratch(nx, ny, nz, cinp, coutp)) {
        set_error("conv3d_wgrad: scratch too small");
        return -1;
    }
    int ngroups, nchunks;
    long long chunk_vox;
    wgrad_plan(V, cinp, coutp, ngroups, nchunks, chunk_vox);
    const int ncoltiles = (27 * cinp + 15) / 16;
    const long long waves = (long long)ngroups * nchunks;
    const dim3 grid((unsigned)((waves + 3) / 4)), block(256);
    switch (coutp / 16) {
    case 1: hipLaunchKernelGGL(k_conv_wgrad<1>, grid, block, 0, s, nx, ny, nz, x, xs, cinp, dy, dys, ncoltiles, ngroups, nchunks, chunk_vox, scratch); break;
    case 2: hipLaunchKernelGGL(k_conv_wgrad<2>, grid, block, 0, s, nx, ny, nz, x, xs, cinp, dy, dys, ncoltiles, ngroups, nchunks, chunk_vox, scratch); break;
    case 3: hipLaunchKernelGGL(k_conv_wgrad<3>, grid, block, 0, s, nx, ny, nz, x, xs, cinp, dy, dys, ncoltiles, ngroups, nchunks, chunk_vox, scratch); break;
    case 4: hipLaunchKernelGGL(k_conv_wgrad<4>, grid, block, 0, s, nx, ny, nz, x, xs, cinp, dy, dys, ncoltiles, ngroups, nchunks, chunk_vox, scratch); break;
    default:
        set_error("conv3d_wgrad: coutp=%d not instantiated (16..64)", coutp);
        return -1;
    }
    DG_CHECK_LAUNCH();
    const size_t nw = (size_t)coutp * 27 * cinp;
    hipLaunchKernelGGL(k_sum_chunks, dim3(grid_for(nw)), dim3(256), 0, s, scratch, nw, nchunks, gw);
    DG_CHECK_LAUNCH();
    if (gb) {
        const long long bchunk = 4096;
        const long long nbc = (V + bchunk - 1) / bchunk;
        float *bp = scratch + (size_t)nchunks * nw;
        hipLaunchKernelGGL(k_bias_part, dim3((unsigned)nbc), dim3(256), 0, s, V, dy, dys, coutp, bchunk, bp);
        DG_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_sum_chunks, dim3(grid_for((size_t)coutp)), dim3(256), 0, s, bp, (size_t)coutp, (int)nbc,
                           gb);
        DG_CHECK_LAUNCH();
    }
    return 0;
}

// ---- resampling, ReLU adjoint, Adam ---------------------------------------------------

// one thread per (coarse voxel, 8-channel group)
__global__ void k_pool2_fwd(int nx, int ny, int nz, int c, const bf16_t *in, int is, bf16_t *out, int os)
{
    const int mx = nx / 2, my = ny / 2, mz = nz / 2, cg = c / 8;
    const size_t n = (size_t)mx * my * mz * cg;
    for (size_t p = (size_t)blockIdx.x * 256 + threadIdx.x; p < n; p += (size_t)gridDim.x * 256) {
        const int g = (int)(p % cg);
        const size_t v = p / cg;
        const int X = (int)(v % mx), Y = (int)((v / mx) % my), Z = (int)(v / ((size_t)mx * my));
        float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int dz = 0; dz < 2; dz++)
            for (int dy = 0; dy < 2; dy++)
                for (int dx = 0; dx < 2; dx++) {
                    const size_t f = ((size_t)(2 * Z + dz) * ny + (2 * Y + dy)) * nx + (2 * X + dx);
                    const bf16_t *q = in + f * is + g * 8;
                    for (int e = 0; e < 8; e++) acc[e] += bf2f(q[e]);
                }
        bf16_t *o = out + v * os + g * 8;
        for (int e = 0; e < 8; e++) o[e] = f2bf(acc[e] * 0.125f);
    }
}

__global__ void k_pool2_bwd_add(int nx, int ny, int nz, int c, const bf16_t *dout, int ds, bf16_t *din, int is)
{
    const int mx = nx / 2, my = ny / 2;
    const size_t n = (size_t)nx * ny * nz * c;
    for (size_t p = (size_t)blockIdx.x * 256 + threadIdx.x; p < n; p += (size_t)gridDim.x * 256) {
        const int ch = (int)(p % c);
        const size_t v = p / c;
        const int X = (int)(v % nx), Y = (int)((v / nx) % ny), Z = (int)(v / ((size_t)nx * ny));
        const size_t pv = ((size_t)(Z / 2) * my + Y / 2) * mx + X / 2;
        bf16_t *d = din + v * is + ch;
        *d = f2bf(bf2f(*d) + 0.125f * bf2f(dout[pv * ds + ch]));
    }
}

__global__ void k_up2_fwd(int nx, int ny, int nz, int c, const bf16_t *in, int is, bf16_t *out, int os)
{
    const int mx = nx / 2, my = ny / 2;
    const int cg = c / 8;
    const size_t n = (size_t)nx * ny * nz * cg;
    for (size_t p = (size_t)blockIdx.x * 256 + threadIdx.x; p < n; p += (size_t)gridDim.x * 256) {
        const int g = (int)(p % cg);
        const size_t v = p / cg;
        const int X = (int)(v % nx), Y = (int)((v / nx) % ny), Z = (int)(v / ((size_t)nx * ny));
        const size_t pv = ((size_t)(Z / 2) * my + Y / 2) * mx + X / 2;
        *reinterpret_cast<uint4 *>(out + v * os + g * 8) = *reinterpret_cast<const uint4 *>(in + pv * is + g * 8);
    }
}

__global__ void k_up2_bwd(int nx, int ny, int nz, int c, const bf16_t *dout, int ds, bf16_t *din, int is)
{
    const int mx = nx / 2, my = ny / 2, mz = nz / 2;
    const size_t n = (size_t)mx * my * mz * c;
    for (size_t p = (size_t)blockIdx.x * 256 + threadIdx.x; p < n; p += (size_t)gridDim.x * 256) {
        const int ch = (int)(p % c);
        const size_t v = p / c;
        const int X = (int)(v % mx), Y = (int)((v / mx) % my), Z = (int)(v / ((size_t)mx * my));
        float acc = 0.f;
        for (int dz = 0; dz < 2; dz++)
            for (int dy = 0; dy < 2; dy++)
                for (int dx = 0; dx < 2; dx++) {
                    const size_t f = ((size_t)(2 * Z + dz) * ny + (2 * Y + dy)) * nx + (2 * X + dx);
                    acc += bf2f(dout[f * ds + ch]);
                }
        din[v * is + ch] = f2bf(acc);
    }
}

__global__ void k_relu_bwd(size_t nvox, int c, const bf16_t *y, int ys, bf16_t *g, int gs)
{
    const size_t n = nvox * c;
    for (size_t p = (size_t)blockIdx.x * 256 + threadIdx.x; p < n; p += (size_t)gridDim.x * 256) {
        const int ch = (int)(p % c);
        const size_t v = p / c;
        if (!(bf2f(y[v * ys + ch]) > 0.f)) g[v * gs + ch] = 0;
    }
}

__global__ void k_adam(size_t n, float *p, const float *g, float *m, float *v, float b1, float b2, float lr_t,
                       float eps_t)
{
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
        const float gq = g[q];
        const float mq = b1 * m[q] + (1.f - b1) * gq;
        const float vq = b2 * v[q] + (1.f - b2) * gq * gq;
        m[q] = mq;
        v[q] = vq;
        p[q] -= lr_t * mq / (sqrtf(vq) + eps_t);
    }
}

static int check_even(int nx, int ny, int nz, int c, const char *who)
{
    if ((nx & 1) || (ny & 1) || (nz & 1) || nx <= 0 || ny <= 0 || nz <= 0 || c <= 0) {
        set_error("%s: extents must be even and positive (%d %d %d), c=%d", who, nx, ny, nz, c);
        return -1;
    }
    return 0;
}

int pool2_fwd(int nx, int ny, int nz, int c, const bf16_t *in, int is, bf16_t *out, int os, hipStream_t s)
{
    if (check_even(nx, ny, nz, c, "pool2_fwd") || (c & 7)) return -1;
    hipLaunchKernelGGL(k_pool2_fwd, dim3(grid_for((size_t)nx * ny * nz / 8 * (c / 8))), dim3(256), 0, s, nx, ny, nz,
                       c, in, is, out, os);
    DG_CHECK_LAUNCH();
    return 0;
}

int pool2_bwd_add(int nx, int ny, int nz, int c, const bf16_t *dout, int ds, bf16_t *din, int is, hipStream_t s)
{
    if (check_even(nx, ny, nz, c, "pool2_bwd_add")) return -1;
    hipLaunchKernelGGL(k_pool2_bwd_add, dim3(grid_for((size_t)nx * ny * nz * c)), dim3(256), 0, s, nx, ny, nz, c,
                       dout, ds, din, is);
    DG_CHECK_LAUNCH();
    return 0;
}

int up2_fwd(int nx, int ny, int nz, int c, const bf16_t *in, int is, bf16_t *out, int os, hipStream_t s)
{
    if (check_even(nx, ny, nz, c, "up2_fwd") || (c & 7) || (is & 7) || (os & 7)) return -1;
    hipLaunchKernelGGL(k_up2_fwd, dim3(grid_for((size_t)nx * ny * nz * (c / 8))), dim3(256), 0, s, nx, ny, nz, c, in,
                       is, out, os);
    DG_CHECK_LAUNCH();
    return 0;
}

int up2_bwd(int nx, int ny, int nz, int c, const bf16_t *dout, int ds, bf16_t *din, int is, hipStream_t s)
{
    if (check_even(nx, ny, nz, c, "up2_bwd")) return -1;
    hipLaunchKernelGGL(k_up2_bwd, dim3(grid_for((size_t)nx * ny * nz / 8 * c)), dim3(256), 0, s, nx, ny, nz, c, dout,
                       ds, din, is);
    DG_CHECK_LAUNCH();
    return 0;
}

int relu_bwd(size_t nvox, int c, const bf16_t *y, int ys, bf16_t *g, int gs, hipStream_t s)
{
    hipLaunchKernelGGL(k_relu_bwd, dim3(grid_for(nvox * c)), dim3(256), 0, s, nvox, c, y, ys, g, gs);
    DG_CHECK_LAUNCH();
    return 0;
}

int adam_update(size_t n, float *p, const float *g, float *m, float *v, float b1, float b2, float lr_t, float eps_t,
                hipStream_t s)
{
    hipLaunchKernelGGL(k_adam, dim3(grid_for(n)), dim3(256), 0, s, n, p, g, m, v, b1, b2, lr_t, eps_t);
    DG_CHECK_LAUNCH();
    return 0;
}

}  // namespace dg
