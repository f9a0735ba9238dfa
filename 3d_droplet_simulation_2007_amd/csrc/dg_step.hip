// dg_step.hip -- HIP kernels of one droplet time step (everything except the
// pressure solver): ghost fills, PLIC-VOF advection, CSF surface tension,
// viscous stress, predictor, convection, projection, k-epsilon, CFL.
//
// FP64 parity semantics: each kernel evaluates its expression in the same
// order as the reference's loop body and the build uses -ffp-contract=off,
// so every operation is one rounded IEEE operation.  Ghost fills run as
// separate phase launches in the reference's x -> y -> z order because later
// phases read values the earlier phases wrote (edge/corner lag, e.g.
// bc_c main.f:4340-4393).

#include <hip/hip_runtime.h>

#include <cmath>

#include "dg_ctx.h"
#include "dg_rb.h"
#include "dg_resident.h"
#include "dg_util.h"

namespace dg {

__device__ __forceinline__ double dmax1(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin1(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double sq(double a) { return a * a; }

// ---------------------------------------------------------------------------
// ghost fills

// bc_scalar (main.f:4298-4329).  Every ghost value is a copy of one interior
// value or 0, and the three passes compose to: ghost(i,j,k) = 0 if i = nx,
// else value at (max(i,2), clamp(j), clamp(k)).  One launch over the faces.
int op_bc_scalar(const Grid &g, hipStream_t s, double *a)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    auto fill = [=] __device__(int i, int j, int k) {
        if (i == nx) { a[g.at(i, j, k)] = 0.0; return; }
        int si = i < 2 ? 2 : i;
        int sj = j < 2 ? 2 : (j > ny - 1 ? ny - 1 : j);
        int sk = k < 2 ? 2 : (k > nz - 1 ? nz - 1 : k);
        a[g.at(i, j, k)] = a[g.at(si, sj, sk)];
    };
    // x faces (all j,k), y faces (all i,k), z faces (all i,j)
    if (launch_face(ny * nz, 2, s, [=] __device__(int p, int f) {
            int j = p % ny + 1, k = p / ny + 1;
            fill(f == 0 ? 1 : nx, j, k);
        }))
        return -1;
    if (launch_face(nx * nz, 2, s, [=] __device__(int p, int f) {
            int i = p % nx + 1, k = p / nx + 1;
            fill(i, f == 0 ? 1 : ny, k);
        }))
        return -1;
    return launch_face(nx * ny, 2, s, [=] __device__(int p, int f) {
        int i = p % nx + 1, j = p / nx + 1;
        fill(i, j, f == 0 ? 1 : nz);
    });
}

// bc_out (main.f:4399-4430)
int op_bc_out(const Grid &g, hipStream_t s, double *a)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (launch_face(ny, nz, s, [=] __device__(int j0, int k0) {
            int j = j0 + 1, k = k0 + 1;
            a[g.at(1, j, k)] = a[g.at(2, j, k)];
            a[g.at(nx, j, k)] = a[g.at(nx - 1, j, k)];
        }))
        return -1;
    if (launch_face(nx, nz - 2, s, [=] __device__(int i0, int k0) {
            int i = i0 + 1, k = k0 + 2;
            a[g.at(i, 1, k)] = a[g.at(i, 2, k)];
            a[g.at(i, ny, k)] = a[g.at(i, ny - 1, k)];
        }))
        return -1;
    return launch_face(nx - 2, ny, s, [=] __device__(int i0, int j0) {
        int i = i0 + 2, j = j0 + 1;
        a[g.at(i, j, 1)] = a[g.at(i, j, 2)];
        a[g.at(i, j, nz)] = a[g.at(i, j, nz - 1)];
    });
}

// bc_resid (main.f:4432-4463)
int op_bc_resid(const Grid &g, hipStream_t s, double *a)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (launch_face(ny, nz, s, [=] __device__(int j0, int k0) {
            int j = j0 + 1, k = k0 + 1;
            a[g.at(1, j, k)] = a[g.at(2, j, k)];
            a[g.at(nx, j, k)] = a[g.at(nx - 1, j, k)];
        }))
        return -1;
    if (launch_face(nx - 2, nz - 2, s, [=] __device__(int i0, int k0) {
            int i = i0 + 2, k = k0 + 2;
            a[g.at(i, 1, k)] = 0.0;
            a[g.at(i, ny, k)] = 0.0;
        }))
        return -1;
    return launch_face(nx - 2, ny, s, [=] __device__(int i0, int j0) {
        int i = i0 + 2, j = j0 + 1;
        a[g.at(i, j, 1)] = 0.0;
        a[g.at(i, j, nz)] = 0.0;
    });
}

// bc_c (main.f:4331-4396); the filter/contact-angle work at main.f:4347-4375
// produces nothing that is used and is not executed.
int op_bc_c(Ctx &c, double *cc)
{
    const Grid g = c.g;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    hipStream_t s = c.stream;
    if (launch_face(ny, nz, s, [=] __device__(int j0, int k0) {
            int j = j0 + 1, k = k0 + 1;
            cc[g.at(2, j, k)] = cc[g.at(1, j, k)];
            cc[g.at(nx, j, k)] = cc[g.at(nx - 1, j, k)];
        }))
        return -1;
    if (launch_face(nx - 2, nz - 2, s, [=] __device__(int i0, int k0) {
            int i = i0 + 2, k = k0 + 2;
            double v = cc[g.at(i, 2, k)];
            if (v > 1.0) v = 1.0;
            if (v < 0.0) v = 0.0;
            cc[g.at(i, 1, k)] = v;
            cc[g.at(i, ny, k)] = cc[g.at(i, ny - 1, k)];
        }))
        return -1;
    return launch_face(nx - 2, ny, s, [=] __device__(int i0, int j0) {
        int i = i0 + 2, j = j0 + 1;
        cc[g.at(i, j, 1)] = cc[g.at(i, j, 2)];
        cc[g.at(i, j, nz)] = cc[g.at(i, j, nz - 1)];
    });
}

// Sequential sums over the x = const planes, k outer and j inner, exactly in
// the order of the reference's outlet loops (main.f:4123-4131): one block
// stages chunks in LDS and lane 0 adds them in order.
__global__ void __launch_bounds__(256) k_ordered_plane_sums(Grid g, const double *A, int ia, const double *B, int ib,
                                                            double *out)
{
    constexpr int CH = 2048;
    __shared__ double sa[CH], sb[CH];
    const int my = g.ny - 2, mz = g.nz - 2;
    const long long tot = (long long)my * mz;
    double accA = 0.0, accB = 0.0;
    for (long long base = 0; base < tot; base += CH) {
        const int cnt = (int)((tot - base) < CH ? (tot - base) : CH);
        for (int t = threadIdx.x; t < cnt; t += 256) {
            long long p = base + t;
            int j = (int)(p % my) + 2, k = (int)(p / my) + 2;
            sa[t] = A[g.at(ia, j, k)];
            sb[t] = B[g.at(ib, j, k)];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int t = 0; t < cnt; t++) {
                accA = accA + sa[t];
                accB = accB + sb[t];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = accA;
        out[1] = accB;
    }
}

static int plane_sums(Ctx &c, const double *A, int ia, const double *B, int ib, double *out)
{
    const Grid g = c.g;
    if (c.ordered_sums) {
        hipLaunchKernelGGL(k_ordered_plane_sums, dim3(1), dim3(256), 0, c.stream, g, A, ia, B, ib, out);
        DG_CHECK_LAUNCH();
        return 0;
    }
    // tree reductions (rounding differs from the sequential order)
    Grid g2 = make_grid(3, g.ny, g.nz);  // iterate j,k interior with i = 2 only
    if (reduce_interior<false>(g2, c.stream, c.red_partial, out, [=] __device__(int, int j, int k) {
            return A[g.at(ia, j, k)];
        }))
        return -1;
    return reduce_interior<false>(g2, c.stream, c.red_partial, out + 1, [=] __device__(int, int j, int k) {
        return B[g.at(ib, j, k)];
    });
}

// bc_v (main.f:4064-4141)
int op_bc_v(Ctx &c)
{
    const Grid g = c.g;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    hipStream_t s = c.stream;
    double *u = c.u, *v = c.v, *w = c.w;
    const double xvel2 = c.s.xvel2;
    if (launch_face(ny, nz, s, [=] __device__(int j0, int k0) {
            int j = j0 + 1, k = k0 + 1;
            double q = ((double)j * 1.0 - 1.5) / ((double)ny * 1.0 - 1.5) - 1.0;
            double val = xvel2 * (-(q * q) + 1.0);
            if (j == 1) {
                double q2 = (2.0 * 1.0 - 1.5) / ((double)ny * 1.0 - 1.5) - 1.0;
                val = -(xvel2 * (-(q2 * q2) + 1.0));
            }
            u[g.at(1, j, k)] = val;
            v[g.at(nx, j, k)] = v[g.at(nx - 1, j, k)];
            w[g.at(nx, j, k)] = w[g.at(nx - 1, j, k)];
        }))
        return -1;
    if (launch_face(nx - 2, nz - 2, s, [=] __device__(int i0, int k0) {
            int i = i0 + 2, k = k0 + 2;
            u[g.at(i, 1, k)] = -u[g.at(i, 2, k)];
            u[g.at(i, ny, k)] = u[g.at(i, ny - 1, k)];
            v[g.at(i, 2, k)] = 0.0;
            v[g.at(i, 1, k)] = 0.0;
            v[g.at(i, ny, k)] = 0.0;
            w[g.at(i, 1, k)] = -w[g.at(i, 2, k)];
            w[g.at(i, ny, k)] = w[g.at(i, ny - 1, k)];
        }))
        return -1;
    if (launch_face(nx - 2, ny, s, [=] __device__(int i0, int j0) {
            int i = i0 + 2, j = j0 + 1;
            u[g.at(i, j, 1)] = u[g.at(i, j, 2)];
            u[g.at(i, j, nz)] = u[g.at(i, j, nz - 1)];
            v[g.at(i, j, 1)] = v[g.at(i, j, 2)];
            v[g.at(i, j, nz)] = v[g.at(i, j, nz - 1)];
            w[g.at(i, j, 2)] = 0.0;
            w[g.at(i, j, nz)] = 0.0;
        }))
        return -1;
    double *sums = c.dscal;
    if (plane_sums(c, u, 1, u, nx - 1, sums)) return -1;
    return launch_face(ny - 2, nz - 2, s, [=] __device__(int j0, int k0) {
        int j = j0 + 2, k = k0 + 2;
        double vout = sums[0] / (double)(ny - 2) / (double)(nz - 2);
        double vout2 = sums[1] / (double)(ny - 2) / (double)(nz - 2);
        u[g.at(2, j, k)] = u[g.at(1, j, k)];
        u[g.at(nx, j, k)] = u[g.at(nx - 1, j, k)] + vout - vout2;
    });
}

// bc_w (main.f:4143-4220)
int op_bc_w(Ctx &c, double *w1, double *w2, double *w3)
{
    const Grid g = c.g;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    hipStream_t s = c.stream;
    const double *u = c.u, *v = c.v, *w = c.w;
    if (launch_face(ny - 2, nz - 2, s, [=] __device__(int j0, int k0) {
            int j = j0 + 2, k = k0 + 2;
            w2[g.at(1, j, k)] = v[g.at(1, j, k)];
            w2[g.at(2, j, k)] = w2[g.at(1, j, k)];
            w2[g.at(nx, j, k)] = v[g.at(nx, j, k)];
            w3[g.at(1, j, k)] = w[g.at(1, j, k)];
            w3[g.at(2, j, k)] = w3[g.at(1, j, k)];
            w3[g.at(nx, j, k)] = w[g.at(nx, j, k)];
        }))
        return -1;
    if (launch_face(nx - 2, nz - 2, s, [=] __device__(int i0, int k0) {
            int i = i0 + 2, k = k0 + 2;
            w1[g.at(i, 1, k)] = -w1[g.at(i, 2, k)];
            w1[g.at(i, ny, k)] = w1[g.at(i, ny - 1, k)];
            w2[g.at(i, 1, k)] = 0.0;
            w2[g.at(i, 2, k)] = 0.0;
            w2[g.at(i, ny, k)] = v[g.at(i, ny, k)];
            w3[g.at(i, 1, k)] = -w3[g.at(i, 2, k)];
            w3[g.at(i, ny, k)] = w3[g.at(i, ny - 1, k)];
        }))
        return -1;
    if (launch_face(nx - 2, ny, s, [=] __device__(int i0, int j0) {
            int i = i0 + 2, j = j0 + 1;
            w1[g.at(i, j, 1)] = w1[g.at(i, j, 2)];
            w1[g.at(i, j, nz)] = w1[g.at(i, j, nz - 1)];
            w2[g.at(i, j, 1)] = w2[g.at(i, j, 2)];
            w2[g.at(i, j, nz)] = w2[g.at(i, j, nz - 1)];
            w3[g.at(i, j, 1)] = w[g.at(i, j, 1)];
            w3[g.at(i, j, 2)] = w3[g.at(i, j, 1)];
            w3[g.at(i, j, nz)] = w[g.at(i, j, nz)];
        }))
        return -1;
    double *sums = c.dscal + 2;
    if (plane_sums(c, u, 1, w1, nx - 1, sums)) return -1;
    return launch_face(ny - 2, nz - 2, s, [=] __device__(int j0, int k0) {
        int j = j0 + 2, k = k0 + 2;
        double vout = sums[0] / (double)(ny - 2) / (double)(nz - 2);
        double vout2 = sums[1] / (double)(ny - 2) / (double)(nz - 2);
        w1[g.at(1, j, k)] = u[g.at(1, j, k)];
        w1[g.at(2, j, k)] = w1[g.at(1, j, k)];
        w1[g.at(nx, j, k)] = w1[g.at(nx - 1, j, k)] + vout - vout2;
    });
}

// bc_tensor (main.f:4222-4296)
int op_bc_tensor(Ctx &c)
{
    const Grid g = c.g;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    hipStream_t s = c.stream;
    double *s11 = c.s11, *s22 = c.s22, *s33 = c.s33, *s12 = c.s12, *s13 = c.s13, *s23 = c.s23;
    if (launch_face(ny - 2, nz - 2, s, [=] __device__(int j0, int k0) {
            int j = j0 + 2, k = k0 + 2;
            s11[g.at(1, j, k)] = 0.0;
            s12[g.at(nx, j, k)] = 0.0;
            s13[g.at(nx, j, k)] = 0.0;
        }))
        return -1;
    if (launch_face(nx, nz - 2, s, [=] __device__(int i0, int k0) {
            int i = i0 + 1, k = k0 + 2;
            s22[g.at(i, 1, k)] = s22[g.at(i, 2, k)];
            s12[g.at(i, ny, k)] = 0.0;
            s23[g.at(i, ny, k)] = 0.0;
        }))
        return -1;
    return launch_face(nx, ny, s, [=] __device__(int i0, int j0) {
        int i = i0 + 1, j = j0 + 1;
        s33[g.at(i, j, 1)] = 0.0;
        s13[g.at(i, j, 2)] = 0.0;
        s13[g.at(i, j, nz)] = 0.0;
        s23[g.at(i, j, 2)] = 0.0;
        s23[g.at(i, j, nz)] = 0.0;
    });
}

int mg_fill(size_t n, hipStream_t s, double *a, double v)
{
    return launch_linear(n, s, [=] __device__(size_t p) { a[p] = v; });
}

// ---------------------------------------------------------------------------
// PLIC plane functions: AL3D (main.f:4562-4634) and FL3D (main.f:4639-4698)

__device__ double al3d(double b1, double b2, double b3, double cc)
{
    const double untier = 1.0 / 3.0;
    double m1, m2, m3, tmp, pr, ch, mm, m12, p, p12, q, teta, cs, v1, v2, v3, cnst, r;
    m1 = dmin1(b1, b2);
    m3 = dmax1(b1, b2);
    m2 = b3;
    if (m2 < m1) {
        tmp = m1; m1 = m2; m2 = tmp;
    } else if (m2 > m3) {
        tmp = m3; m3 = m2; m2 = tmp;
    }
    m12 = m1 + m2;
    pr = dmax1(6.0 * m1 * m2 * m3, 1.0e-50);
    v1 = m1 * m1 * m1 / pr;
    v2 = v1 + 0.5 * (m2 - m1) / m3;
    if (m3 < m12) {
        mm = m3;
        v3 = (m3 * m3 * (3.0 * m12 - m3) + m1 * m1 * (m1 - 3.0 * m3) + m2 * m2 * (m2 - 3.0 * m3)) / pr;
    } else {
        mm = m12;
        v3 = 0.5 * mm / m3;
    }
    ch = dmin1(cc, 1.0 - cc);
    if (ch < v1) {
        r = pow(pr * ch, untier);
    } else if (ch < v2) {
        r = 0.5 * (m1 + sqrt(m1 * m1 + 8.0 * m2 * m3 * (ch - v1)));
    } else if (ch < v3) {
        p = 2.0 * m1 * m2;
        q = 1.5 * m1 * m2 * (m12 - 2.0 * m3 * ch);
        p12 = sqrt(p);
        cnst = q / (p * p12);
        if (cnst > 1.0) cnst = 1.0;
        if (cnst < -1.0) cnst = -1.0;
        teta = acos(cnst) / 3.0;
        cs = cos(teta);
        r = p12 * (sqrt(3.0 * (1.0 - cs * cs)) - cs) + m12;
    } else if (mm == m12) {
        r = m3 * ch + 0.5 * mm;
    } else {
        p = m1 * (m2 + m3) + m2 * m3 - 0.25;
        q = 1.5 * m1 * m2 * m3 * (0.5 - ch);
        p12 = sqrt(p);
        cnst = q / (p * p12);
        if (cnst > 1.0) cnst = 1.0;
        if (cnst < -1.0) cnst = -1.0;
        teta = acos(cnst) / 3.0;
        cs = cos(teta);
        r = p12 * (sqrt(3.0 * (1.0 - cs * cs)) - cs) + 0.5;
    }
    if (cc > 0.5) r = 1.0 - r;
    return r;
}

__device__ double fl3d(double m1, double m2, double m3, double alpha, double r0, double dr0)
{
    double al, al0, n1, n2, n3, b1, b2, b3, b12, bm, tmp, pr;
    al = alpha - m1 * r0;
    al = al + dmax1(0.0, -m1 * dr0) + dmax1(0.0, -m2) + dmax1(0.0, -m3);
    tmp = fabs(m1) * dr0 + fabs(m2) + fabs(m3);
    n1 = fabs(m1) / tmp;
    n2 = fabs(m2) / tmp;
    n3 = fabs(m3) / tmp;
    al = dmax1(0.0, dmin1(1.0, al / tmp));
    al0 = dmin1(al, 1.0 - al);
    b1 = dmin1(n1 * dr0, n2);
    b3 = dmax1(n1 * dr0, n2);
    b2 = n3;
    if (b2 < b1) {
        tmp = b1; b1 = b2; b2 = tmp;
    } else if (b2 > b3) {
        tmp = b3; b3 = b2; b2 = tmp;
    }
    b12 = b1 + b2;
    bm = dmin1(b12, b3);
    pr = dmax1(6.0 * b1 * b2 * b3, 1.0e-50);
    if (al0 < b1) {
        tmp = al0 * al0 * al0 / pr;
    } else if (al0 < b2) {
        tmp = 0.5 * al0 * (al0 - b1) / (b2 * b3) + b1 * b1 * b1 / pr;
    } else if (al0 < bm) {
        tmp = (al0 * al0 * (3.0 * b12 - al0) + b1 * b1 * (b1 - 3.0 * al0) + b2 * b2 * (b2 - 3.0 * al0)) / pr;
    } else if (bm == b12) {
        tmp = (al0 - 0.5 * bm) / b3;
    } else {
        tmp = (al0 * al0 * (3.0 - 2.0 * al0) + b1 * b1 * (b1 - 3.0 * al0) + b2 * b2 * (b2 - 3.0 * al0) +
               b3 * b3 * (b3 - 3.0 * al0)) / pr;
    }
    if (al <= 0.5) return tmp * dr0;
    return (1.0 - tmp) * dr0;
}

int plic_al3d_batch(int n, hipStream_t s, const double *b1, const double *b2, const double *b3, const double *cc,
                    double *out)
{
    return launch_linear((size_t)n, s, [=] __device__(size_t p) { out[p] = al3d(b1[p], b2[p], b3[p], cc[p]); });
}

int plic_fl3d_batch(int n, hipStream_t s, const double *m1, const double *m2, const double *m3, const double *al,
                    const double *r0, const double *dr0, double *out)
{
    return launch_linear((size_t)n, s,
                         [=] __device__(size_t p) { out[p] = fl3d(m1[p], m2[p], m3[p], al[p], r0[p], dr0[p]); });
}

// ---------------------------------------------------------------------------
// VOF: swp (main.f:848-1000) and VOFmethod (main.f:795-824)

// The fluxes of cell o in sweep direction d (swp main.f:848-1000): f1 leaves
// through the lower face, f2 stays, f3 leaves through the upper face, with
// a1 = us(o) tauh, a2 = us(o + sd) tauh.  want selects the fluxes formed
// (1: f1, 2: f2, 4: f3); each is formed exactly as when all three are, and
// the Youngs normal and the plane are skipped when no wanted flux needs them.
// c0 is cc(o).
__device__ inline void swp_fluxes(const double *cc, const Grid g, int o, double c0, double a1, double a2, int d,
                                  int guard, int want, double &f1, double &f2, double &f3)
{
    const double tiny = DG_F4(1.e-30);
    f1 = 0.0;
    f2 = 0.0;
    f3 = 0.0;
#define C(di, dj, dk) cc[o + (di) + g.sy * (dj) + g.sz * (dk)]
    if (c0 == 1.0) {
        f1 = dmax1(-a1, 0.0);
        f2 = 1.0 - dmax1(a1, 0.0) + dmin1(a2, 0.0);
        f3 = dmax1(a2, 0.0);
    } else if (c0 > 0.0) {
        const bool w1 = (want & 1) && a1 < 0.0, w3 = (want & 4) && a2 > 0.0;
        if (!(w1 || w3 || (want & 2))) return;
        double mm1, mm2, mx, my, mz;
        mm1 = C(-1, -1, -1) + C(-1, -1, 1) + C(-1, 1, -1) + C(-1, 1, 1) +
              2.0 * (C(-1, -1, 0) + C(-1, 1, 0) + C(-1, 0, -1) + C(-1, 0, 1)) + 4.0 * C(-1, 0, 0);
        mm2 = C(1, -1, -1) + C(1, -1, 1) + C(1, 1, -1) + C(1, 1, 1) +
              2.0 * (C(1, -1, 0) + C(1, 1, 0) + C(1, 0, -1) + C(1, 0, 1)) + 4.0 * C(1, 0, 0);
        mx = mm1 - mm2;
        mm1 = C(-1, -1, -1) + C(-1, -1, 1) + C(1, -1, -1) + C(1, -1, 1) +
              2.0 * (C(-1, -1, 0) + C(1, -1, 0) + C(0, -1, -1) + C(0, -1, 1)) + 4.0 * C(0, -1, 0);
        mm2 = C(-1, 1, -1) + C(-1, 1, 1) + C(1, 1, -1) + C(1, 1, 1) +
              2.0 * (C(-1, 1, 0) + C(1, 1, 0) + C(0, 1, -1) + C(0, 1, 1)) + 4.0 * C(0, 1, 0);
        my = mm1 - mm2;
        mm1 = C(-1, -1, -1) + C(-1, 1, -1) + C(1, -1, -1) + C(1, 1, -1) +
              2.0 * (C(-1, 0, -1) + C(1, 0, -1) + C(0, -1, -1) + C(0, 1, -1)) + 4.0 * C(0, 0, -1);
        mm2 = C(-1, -1, 1) + C(-1, 1, 1) + C(1, -1, 1) + C(1, 1, 1) +
              2.0 * (C(-1, 0, 1) + C(1, 0, 1) + C(0, -1, 1) + C(0, 1, 1)) + 4.0 * C(0, 0, 1);
        mz = mm1 - mm2;
        double invx = 1.0, invy = 1.0, invz = 1.0;
        if (mx < 0.0) { mx = -mx; invx = -1.0; }
        if (my < 0.0) { my = -my; invy = -1.0; }
        if (mz < 0.0) { mz = -mz; invz = -1.0; }
        if (guard && mx + my + mz == 0.0) {
            if (d == 1) mx = 1.0;
            else if (d == 2) my = 1.0;
            else mz = 1.0;
        }
        mm2 = mx + my + mz + tiny;
        mx = mx / mm2;
        my = my / mm2;
        mz = mz / mm2;
        double alpha = al3d(mx, my, mz, c0);
        mx = invx * mx;
        my = invy * my;
        mz = invz * mz;
        alpha = alpha + dmin1(0.0, mx) + dmin1(0.0, my) + dmin1(0.0, mz);
        mm1 = dmax1(a1, 0.0);
        mm2 = 1.0 - mm1 + dmin1(0.0, a2);
        // the advected direction goes first in FL3D's argument list
        double md, mp, mq;
        if (d == 1) { md = mx; mp = my; mq = mz; }
        else if (d == 2) { md = my; mp = mz; mq = mx; }
        else { md = mz; mp = mx; mq = my; }
        md = md / (1.0 - a1 + a2);
        alpha = alpha + md * a1;
        if (w1) f1 = fl3d(md, mp, mq, alpha, a1, -a1);
        if (w3) f3 = fl3d(md, mp, mq, alpha, 1.0, a2);
        if (want & 2) f2 = fl3d(md, mp, mq, alpha, mm1, mm2);
    }
#undef C
}

// One sweep fused into one pass (the time step's path): cell o's new volume
// fraction vof3(o - sd) + vof2(o) + vof1(o + sd) is formed from the fluxes of
// its two neighbours along d, recomputed here, instead of three flux arrays
// written and read back.  Only the interface cells whose neighbour's
// velocity points at them do PLIC work twice, so the pass moves cc, the
// velocity and the new cc (24 B per cell against 72).  Out of place: every
// thread reads the old cc; ghost cells are copied.  The neighbour fluxes
// across the domain's ends are bc_flux's zeros (main.f:4496-4539).  One
// cell per thread in the dispatcher's own tile order, which deals the few
// tiles with interface cells (and their PLIC work) over all XCDs: walking
// columns in XCD-contiguous tiles fetched 30% less but took 1.7x as long.
template <int D>
__global__ void __launch_bounds__(256) k_swp_fused(Grid g, const double *us, const double *cc, double *ccn,
                                                   double tauh, int guard)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1;
    const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) + 1;
    const int k = (int)blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const int o = g.at(i, j, k);
    if (i == 1 || i == g.nx || j == 1 || j == g.ny || k == 1 || k == g.nz) {
        ccn[o] = cc[o];
        return;
    }
    const int sd = D == 1 ? 1 : (D == 2 ? g.sy : g.sz);
    const int id = D == 1 ? i : (D == 2 ? j : k), nd = D == 1 ? g.nx : (D == 2 ? g.ny : g.nz);
    const double u0 = us[o] * tauh, up = us[o + sd] * tauh;
    double f1, f2, f3, fb = 0.0, fa = 0.0, t1, t2;
    swp_fluxes(cc, g, o, cc[o], u0, up, D, guard, 2, f1, f2, f3);
    if (id > 2) swp_fluxes(cc, g, o - sd, cc[o - sd], us[o - sd] * tauh, u0, D, guard, 4, t1, t2, fb);
    if (id < nd - 1) swp_fluxes(cc, g, o + sd, cc[o + sd], up, us[o + 2 * sd] * tauh, D, guard, 1, fa, t1, t2);
    const double v = fb + f2 + fa;
    ccn[o] = dmax1(0.0, dmin1(1.0, v));
}

// bc_flux's zeros (main.f:4496-4539) on the ghost faces of the flux arrays
static int bc_flux(const Grid &g, hipStream_t s, double *vof1, double *vof3, int d)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (d == 1)
        return launch_face(ny - 2, nz - 2, s, [=] __device__(int j0, int k0) {
            vof3[g.at(1, j0 + 2, k0 + 2)] = 0.0;
            vof1[g.at(nx, j0 + 2, k0 + 2)] = 0.0;
        });
    if (d == 2)
        return launch_face(nx - 2, nz - 2, s, [=] __device__(int i0, int k0) {
            vof3[g.at(i0 + 2, 1, k0 + 2)] = 0.0;
            vof1[g.at(i0 + 2, ny, k0 + 2)] = 0.0;
        });
    return launch_face(nx - 2, ny - 2, s, [=] __device__(int i0, int j0) {
        vof3[g.at(i0 + 2, j0 + 2, 1)] = 0.0;
        vof1[g.at(i0 + 2, j0 + 2, nz)] = 0.0;
    });
}

static int swp(Ctx &c, const double *us, double *cc, double *vof1, double *vof2, double *vof3, int d)
{
    const Grid g = c.g;
    const double tauh = c.s.tauh;
    const int sd = d == 1 ? 1 : (d == 2 ? g.sy : g.sz);
    const int guard = c.vof_guard;
    if (launch_interior_rr(g, c.stream, [=] __device__(int i, int j, int k) {
            const int o = g.at(i, j, k);
            double f1, f2, f3;
            swp_fluxes(cc, g, o, cc[o], us[o] * tauh, us[o + sd] * tauh, d, guard, 7, f1, f2, f3);
            vof1[o] = f1;
            vof2[o] = f2;
            vof3[o] = f3;
        }))
        return -1;
    if (bc_flux(g, c.stream, vof1, vof3, d)) return -1;
    if (launch_interior(g, c.stream, [=] __device__(int i, int j, int k) {
            const int o = g.at(i, j, k);
            double v = vof3[o - sd] + vof2[o] + vof1[o + sd];
            cc[o] = dmax1(0.0, dmin1(1.0, v));
        }))
        return -1;
    return op_bc_c(c, cc);
}

// swp for the time step: one fused pass into csf's ccf array (free until
// csf rewrites it whole), whose pointer then trades places with cc.  The
// flux arrays keep the ghost zeros of bc_flux; their interiors are not
// written (the step rewrites them before reading them: kepsmodel, csf's
// stress copies, solvelpres's predictor and abcei).
static int swp_step(Ctx &c, const double *us, double *vof1, double *vof3, int d)
{
    const Grid g = c.g;
    Level &L = c.lev[c.ng - 1];
    const dim3 grid((unsigned)((g.nx + BX - 1) / BX), (unsigned)((g.ny + BY - 1) / BY), (unsigned)g.nz);
    const double tauh = c.s.tauh;
    if (d == 1)
        hipLaunchKernelGGL(k_swp_fused<1>, grid, interior_block(), 0, c.stream, g, us, L.cc, c.ccf, tauh, c.vof_guard);
    else if (d == 2)
        hipLaunchKernelGGL(k_swp_fused<2>, grid, interior_block(), 0, c.stream, g, us, L.cc, c.ccf, tauh, c.vof_guard);
    else
        hipLaunchKernelGGL(k_swp_fused<3>, grid, interior_block(), 0, c.stream, g, us, L.cc, c.ccf, tauh, c.vof_guard);
    DG_CHECK_LAUNCH();
    if (bc_flux(g, c.stream, vof1, vof3, d)) return -1;
    double *t = L.cc;
    L.cc = c.ccf;
    c.ccf = t;
    return op_bc_c(c, L.cc);
}

int op_swp(Ctx &c, const double *us, double *cc, double *vof1, double *vof2, double *vof3, int d)
{
    if (d < 1 || d > 3) {
        set_error("swp: direction %d is not 1, 2 or 3", d);
        return -1;
    }
    return swp(c, us, cc, vof1, vof2, vof3, d);
}

// fturbdif (main.f:1239-1267): in-place update with k innermost, then j, so
// cell (j,k) sees the new (j-1,k) and (j,k-1).  The stencil has no x
// neighbours: one block per x plane sweeps the (j,k) anti-diagonals.
__global__ void __launch_bounds__(256) k_fturbdif(Grid g, double *cc, const double *mut, const double *rh, double ftr)
{
    const int i = (int)blockIdx.x + 2;
    const int my = g.ny - 2, mz = g.nz - 2;
    for (int dsum = 0; dsum <= my + mz - 2; dsum++) {
        // cells with (j-2)+(k-2) = dsum
        int jlo = dsum - (mz - 1) > 0 ? dsum - (mz - 1) : 0;
        int jhi = dsum < my - 1 ? dsum : my - 1;
        for (int jj = jlo + (int)threadIdx.x; jj <= jhi; jj += 256) {
            int j = jj + 2, k = dsum - jj + 2;
            int o = g.at(i, j, k);
            double m0 = mut[o] / rh[o];
            double c0 = cc[o];
            cc[o] = c0 + ftr * ((m0 + mut[o + g.sy] / rh[o + g.sy]) * (cc[o + g.sy] - c0) -
                                (m0 + mut[o - g.sy] / rh[o - g.sy]) * (c0 - cc[o - g.sy]) +
                                (m0 + mut[o + g.sz] / rh[o + g.sz]) * (cc[o + g.sz] - c0) -
                                (m0 + mut[o - g.sz] / rh[o - g.sz]) * (c0 - cc[o - g.sz]));
        }
        __syncthreads();
    }
}

int op_calcrho(Ctx &c)
{
    const double rho1 = c.s.rho1, rho2 = c.s.rho2, mu1 = c.s.mu1, mu2 = c.s.mu2;
    const double *cc = c.lev[c.ng - 1].cc;
    double *rh = c.rhon, *mn = c.mun;
    return launch_linear(c.g.n(), c.stream, [=] __device__(size_t p) {
        rh[p] = rho2 + cc[p] * (rho1 - rho2);
        mn[p] = mu2 + cc[p] * (mu1 - mu2);
    });
}

int op_vofmethod(Ctx &c, int t, bool step)
{
    (void)t;
    Level &L = c.lev[c.ng - 1];
    c.s.tswap = c.s.tswap + 1;
    if (c.s.tswap > 3) c.s.tswap = 1;
    int order[3];
    if (c.s.tswap % 3 == 0) { order[0] = 3; order[1] = 1; order[2] = 2; }
    else if (c.s.tswap % 2 == 0) { order[0] = 2; order[1] = 3; order[2] = 1; }
    else { order[0] = 1; order[1] = 2; order[2] = 3; }
    for (int q = 0; q < 3; q++) {
        const double *us = order[q] == 1 ? c.u : (order[q] == 2 ? c.v : c.w);
        if (step ? swp_step(c, us, L.ei, L.rhs, order[q]) : swp(c, us, L.cc, L.ei, L.res, L.rhs, order[q])) return -1;
    }
    if (op_bc_c(c, L.cc)) return -1;
    if (c.s.iturb != 0) {
        double ftr = DG_F4(1.0);
        ftr = ftr * (double)c.s.tauh2 * 1.0 / 2.0;
        hipLaunchKernelGGL(k_fturbdif, dim3((unsigned)(c.g.nx - 2)), dim3(256), 0, c.stream, c.g, L.cc, c.mut, c.rhon,
                           ftr);
        DG_CHECK_LAUNCH();
        if (op_bc_c(c, L.cc)) return -1;
    }
    return op_calcrho(c);
}

// ---------------------------------------------------------------------------
// k-epsilon transport, kepsmodel (main.f:1888-2026), iturb = 3

int op_kepsmodel(Ctx &c, int t)
{
    if (c.s.iturb == 4 && alloc_reynolds(c)) return -1;
    // the production term from the Reynolds stresses (iturb = 4, after the
    // first step) or from the eddy viscosity (main.f:1955-1969)
    const bool rey = c.s.iturb == 4 && t != 0;
    const double *u2p = c.u2p, *v2p = c.v2p, *w2p = c.w2p, *uvp = c.uvp, *uwp = c.uwp, *vwp = c.vwp;
    const Grid g = c.g;
    Level &L = c.lev[c.ng - 1];
    double *kn = L.u, *en = L.res;
    const double *u = c.u, *v = c.v, *w = c.w, *kt = c.ktur, *ep = c.eps, *mut = c.mut, *mun = c.mun, *rh = c.rhon;
    const double h = c.s.h, tau = c.s.tau;
    const double h2 = h * h, h22 = h2 * 2.0;
    const int sy = g.sy, sz = g.sz;
    if (launch_interior(g, c.stream, [=] __device__(int i, int j, int k) {
            const double sigk = 1.0, sige = 1.3, c1 = 1.44, c2 = 1.92;
            const int o = g.at(i, j, k);
#define U(di, dj, dk) u[o + (di) + sy * (dj) + sz * (dk)]
#define V(di, dj, dk) v[o + (di) + sy * (dj) + sz * (dk)]
#define W(di, dj, dk) w[o + (di) + sy * (dj) + sz * (dk)]
#define F(a, di, dj, dk) a[o + (di) + sy * (dj) + sz * (dk)]
            double conv, diff, prod, cip, cim, cjp, cjm, ckp, ckm;
            conv = (+dmax1(U(1, 0, 0), 0.0) * F(kt, 0, 0, 0) + dmin1(U(1, 0, 0), 0.0) * F(kt, 1, 0, 0) -
                    dmax1(U(0, 0, 0), 0.0) * F(kt, -1, 0, 0) - dmin1(U(0, 0, 0), 0.0) * F(kt, 0, 0, 0) +
                    dmax1(V(0, 1, 0), 0.0) * F(kt, 0, 0, 0) + dmin1(V(0, 1, 0), 0.0) * F(kt, 0, 1, 0) -
                    dmax1(V(0, 0, 0), 0.0) * F(kt, 0, -1, 0) - dmin1(V(0, 0, 0), 0.0) * F(kt, 0, 0, 0) +
                    dmax1(W(0, 0, 1), 0.0) * F(kt, 0, 0, 0) + dmin1(W(0, 0, 1), 0.0) * F(kt, 0, 0, 1) -
                    dmax1(W(0, 0, 0), 0.0) * F(kt, 0, 0, -1) - dmin1(W(0, 0, 0), 0.0) * F(kt, 0, 0, 0)) /
                   h;
            cip = F(mun, 1, 0, 0) + F(mun, 0, 0, 0) + (F(mut, 0, 0, 0) + F(mut, 1, 0, 0)) / sigk;
            cim = F(mun, -1, 0, 0) + F(mun, 0, 0, 0) + (F(mut, 0, 0, 0) + F(mut, -1, 0, 0)) / sigk;
            cjp = F(mun, 0, 1, 0) + F(mun, 0, 0, 0) + (F(mut, 0, 0, 0) + F(mut, 0, 1, 0)) / sigk;
            cjm = F(mun, 0, -1, 0) + F(mun, 0, 0, 0) + (F(mut, 0, 0, 0) + F(mut, 0, -1, 0)) / sigk;
            ckp = F(mun, 0, 0, 1) + F(mun, 0, 0, 0) + (F(mut, 0, 0, 0) + F(mut, 0, 0, 1)) / sigk;
            ckm = F(mun, 0, 0, -1) + F(mun, 0, 0, 0) + (F(mut, 0, 0, 0) + F(mut, 0, 0, -1)) / sigk;
            diff = 1.0 / (F(rh, 0, 0, 0) * h22) *
                   (cip * (F(kt, 1, 0, 0) - F(kt, 0, 0, 0)) - cim * (F(kt, 0, 0, 0) - F(kt, -1, 0, 0)) +
                    cjp * (F(kt, 0, 1, 0) - F(kt, 0, 0, 0)) - cjm * (F(kt, 0, 0, 0) - F(kt, 0, -1, 0)) +
                    ckp * (F(kt, 0, 0, 1) - F(kt, 0, 0, 0)) - ckm * (F(kt, 0, 0, 0) - F(kt, 0, 0, -1)));
            double dudx = U(1, 0, 0) - U(0, 0, 0);
            double dudy = 0.25 * (U(0, 0, 0) - U(0, -1, 0) + U(1, 0, 0) - U(1, -1, 0) + U(0, 1, 0) - U(0, 0, 0) +
                                  U(1, 1, 0) - U(1, 0, 0));
            double dudz = 0.25 * (U(0, 0, 0) - U(0, 0, -1) + U(1, 0, 0) - U(1, 0, -1) + U(0, 0, 1) - U(0, 0, 0) +
                                  U(1, 0, 1) - U(1, 0, 0));
            double dvdx = 0.25 * (V(0, 0, 0) - V(-1, 0, 0) + V(1, 0, 0) - V(0, 0, 0) + V(0, 1, 0) - V(-1, 1, 0) +
                                  V(1, 1, 0) - V(0, 1, 0));
            double dvdy = V(0, 1, 0) - V(0, 0, 0);
            double dvdz = 0.25 * (V(0, 0, 0) - V(0, 0, -1) + V(0, 0, 1) - V(0, 0, 0) + V(0, 1, 0) - V(0, 1, -1) +
                                  V(0, 1, 1) - V(0, 1, 0));
            double dwdx = 0.25 * (W(0, 0, 0) - W(-1, 0, 0) + W(1, 0, 0) - W(0, 0, 0) + W(0, 0, 1) - W(-1, 0, 1) +
                                  W(1, 0, 1) - W(0, 0, 1));
            double dwdy = 0.25 * (W(0, 0, 0) - W(0, -1, 0) + W(0, 0, 1) - W(0, -1, 1) + W(0, 1, 0) - W(0, 0, 0) +
                                  W(0, 1, 1) - W(0, 0, 1));
            double dwdz = W(0, 0, 1) - W(0, 0, 0);
            if (!rey)
                prod = F(mut, 0, 0, 0) / F(rh, 0, 0, 0) / h2 *
                       (2.0 * (dudx * dudx + dvdy * dvdy + dwdz * dwdz + dvdx * dudy + dwdx * dudz + dvdz * dwdy) +
                        dudy * dudy + dudz * dudz + dvdx * dvdx + dvdz * dvdz + dwdx * dwdx + dwdy * dwdy);
            else
                prod = -(u2p[o] * dudx + v2p[o] * dvdy + w2p[o] * dwdz + uvp[o] * (dudy + dvdx) +
                         uwp[o] * (dudz + dwdx) + vwp[o] * (dvdz + dwdy)) /
                       h;
            const double knew = F(kt, 0, 0, 0) + tau * (-conv + diff + prod - F(ep, 0, 0, 0));
            conv = (+dmax1(U(1, 0, 0), 0.0) * F(ep, 0, 0, 0) + dmin1(U(1, 0, 0), 0.0) * F(ep, 1, 0, 0) -
                    dmax1(U(0, 0, 0), 0.0) * F(ep, -1, 0, 0) - dmin1(U(0, 0, 0), 0.0) * F(ep, 0, 0, 0) +
                    dmax1(V(0, 1, 0), 0.0) * F(ep, 0, 0, 0) + dmin1(V(0, 1, 0), 0.0) * F(ep, 0, 1, 0) -
                    dmax1(V(0, 0, 0), 0.0) * F(ep, 0, -1, 0) - dmin1(V(0, 0, 0), 0.0) * F(ep, 0, 0, 0) +
                    dmax1(W(0, 0, 1), 0.0) * F(ep, 0, 0, 0) + dmin1(W(0, 0, 1), 0.0) * F(ep, 0, 0, 1) -
                    dmax1(W(0, 0, 0), 0.0) * F(ep, 0, 0, -1) - dmin1(W(0, 0, 0), 0.0) * F(ep, 0, 0, 0)) /
                   h;
            cip = F(mun, 1, 0, 0) + F(mun, 0, 0, 0) + (F(mut, 0, 0, 0) + F(mut, 1, 0, 0)) / sige;
            cim = F(mun, -1, 0, 0) + F(mun, 0, 0, 0) + (F(mut, 0, 0, 0) + F(mut, -1, 0, 0)) / sige;
            cjp = F(mun, 0, 1, 0) + F(mun, 0, 0, 0) + (F(mut, 0, 0, 0) + F(mut, 0, 1, 0)) / sige;
            cjm = F(mun, 0, -1, 0) + F(mun, 0, 0, 0) + (F(mut, 0, 0, 0) + F(mut, 0, -1, 0)) / sige;
            ckp = F(mun, 0, 0, 1) + F(mun, 0, 0, 0) + (F(mut, 0, 0, 0) + F(mut, 0, 0, 1)) / sige;
            ckm = F(mun, 0, 0, -1) + F(mun, 0, 0, 0) + (F(mut, 0, 0, 0) + F(mut, 0, 0, -1)) / sige;
            diff = 1.0 / (F(rh, 0, 0, 0) * h22) *
                   (cip * (F(ep, 1, 0, 0) - F(ep, 0, 0, 0)) - cim * (F(ep, 0, 0, 0) - F(ep, -1, 0, 0)) +
                    cjp * (F(ep, 0, 1, 0) - F(ep, 0, 0, 0)) - cjm * (F(ep, 0, 0, 0) - F(ep, 0, -1, 0)) +
                    ckp * (F(ep, 0, 0, 1) - F(ep, 0, 0, 0)) - ckm * (F(ep, 0, 0, 0) - F(ep, 0, 0, -1)));
            prod = prod * c1 * F(ep, 0, 0, 0) / F(kt, 0, 0, 0);
            const double enew = F(ep, 0, 0, 0) +
                    tau * (-conv + diff + prod - c2 * F(ep, 0, 0, 0) * F(ep, 0, 0, 0) / F(kt, 0, 0, 0));
            kn[o] = knew;
            en[o] = enew;
#undef U
#undef V
#undef W
#undef F
        }))
        return -1;
    double *ktur = c.ktur, *epsa = c.eps, *muta = c.mut;
    const double *cmu = c.cmu;
    if (launch_interior(g, c.stream, [=] __device__(int i, int j, int k) {
            const int o = g.at(i, j, k);
            const double kv = kn[o], ev = en[o];
            ktur[o] = kv;
            epsa[o] = ev;
            muta[o] = cmu[o] * rh[o] * kv * kv / ev;
        }))
        return -1;
    if (op_bc_scalar(g, c.stream, c.ktur)) return -1;
    if (op_bc_scalar(g, c.stream, c.eps)) return -1;
    if (op_bc_scalar(g, c.stream, c.mut)) return -1;
    return c.s.iturb == 4 ? op_uv2p(c) : 0;  // main.f:2023
}

// eddyvisc (main.f:1828-1885), iturb = 1, 2: u filtered into the finest
// pressure slot, which the reference passes as ufilt (main.f:35); the
// mixing length of each x station from the filtered profile through
// (j1, k1); mut from it and the y and z shear of u.
//
// The station scan carries j2 from one station to the next when the second
// branch finds no match (main.f:1855-1861), so one workgroup finds each
// station's first matching j in parallel and one thread then walks the
// stations in order.  j2 is uninitialised in the reference before that
// branch's first station; it starts at 0 here.
__global__ void __launch_bounds__(256) k_eddyvisc_lmix(Grid g, const double *uf, int j1, int k1, int jet,
                                                       double xvel2, double alfa_h, double *lmix, int *found)
{
    for (int i = (int)threadIdx.x + 1; i <= g.nx; i += (int)blockDim.x) {
        const double centre = uf[g.at(i, j1, k1)];
        const double umax = jet ? centre : xvel2, umin = jet ? xvel2 : centre;
        const double lim = DG_F4(0.1) * (umax - umin);
        int f = 0;
        for (int j = 2; j <= g.ny - 1; j++) {
            const double d = uf[g.at(i, j, k1)] - umin;
            if (jet ? d > lim : d <= lim) {
                f = j;
                break;
            }
        }
        found[i - 1] = f;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int j2 = 0;
        for (int i = 1; i <= g.nx; i++) {
            if (jet) j2 = found[i - 1];
            else if (found[i - 1]) j2 = found[i - 1];
            if (j2 == 2) j2 = j1;
            lmix[i - 1] = alfa_h * (double)abs(j2 - j1);
        }
    }
}

int op_eddyvisc(Ctx &c)
{
    const Grid g = c.g;
    double *uf = c.lev[c.ng - 1].u;
    const double *u = c.u;
    const int sy = g.sy, sz = g.sz;
    // filter (main.f:1271-1284)
    if (launch_interior(g, c.stream, [=] __device__(int i, int j, int k) {
            const int o = g.at(i, j, k);
            uf[o] = 0.25 * u[o] + 0.125 * (u[o - sz] + u[o + sz] + u[o - sy] + u[o + sy]) +
                    0.0625 * (u[o - sy - sz] + u[o - sy + sz] + u[o + sy - sz] + u[o + sy + sz]);
        }))
        return -1;
    const double h = c.s.h;
    const int j1 = (int)(c.s.y0 / h), k1 = (int)(c.s.z0 / h);
    if (j1 < 1 || j1 > g.ny || k1 < 1 || k1 > g.nz) {
        set_error("eddyvisc: centre station (j1, k1) = (%d, %d) is outside the %d x %d x %d grid (main.f:1841-1842)",
                  j1, k1, g.nx, g.ny, g.nz);
        return -1;
    }
    // the stations' first matches, as ints in the second half of lmix's allocation
    int *found = reinterpret_cast<int *>(c.lmix + g.nx);
    hipLaunchKernelGGL(k_eddyvisc_lmix, dim3(1), dim3(256), 0, c.stream, g, uf, j1, k1, c.s.xvel1 > c.s.xvel2 ? 1 : 0,
                       c.s.xvel2, DG_F4(0.08) * h, c.lmix, found);
    DG_CHECK_LAUNCH();
    const double *rh = c.rhon, *lmix = c.lmix;
    double *mut = c.mut;
    const double cmu1 = DG_F4(1.0);
    if (launch_interior(g, c.stream, [=] __device__(int i, int j, int k) {
            const int o = g.at(i, j, k);
            const double dudy = 0.25 * (u[o] - u[o - sy] + u[o + 1] - u[o + 1 - sy] + u[o + sy] - u[o] +
                                        u[o + 1 + sy] - u[o + 1]) /
                                h;
            const double dudz = 0.25 * (u[o] - u[o - sz] + u[o + 1] - u[o + 1 - sz] + u[o + sz] - u[o] +
                                        u[o + 1 + sz] - u[o + 1]) /
                                h;
            const double lm = lmix[i - 1] + lmix[i];
            mut[o] = cmu1 * rh[o] * 0.25 * lm * lm * sqrt(dudy * dudy + dudz * dudz);
        }))
        return -1;
    return op_bc_scalar(g, c.stream, mut);
}

// uv2p (main.f:2032-2155), iturb = 4: Shih-Zhu-Lumley eddy viscosity and
// Reynolds stresses.  The reference's REAL constants are rounded to single
// precision: 1./3., 2./3., 1.E-30 and sqrt(6.).
int op_uv2p(Ctx &c)
{
    if (alloc_reynolds(c)) return -1;
    const Grid g = c.g;
    const double *u = c.u, *v = c.v, *w = c.w, *kt = c.ktur, *ep = c.eps, *rh = c.rhon;
    double *cmu = c.cmu, *mut = c.mut, *u2p = c.u2p, *v2p = c.v2p, *w2p = c.w2p, *uvp = c.uvp, *uwp = c.uwp,
           *vwp = c.vwp;
    const double h = c.s.h;
    const double third = (double)(1.0f / 3.0f), twothird = (double)(2.0f / 3.0f), tiny = DG_F4(1.E-30);
    const double sixsq = (double)sqrtf(6.0f);
    const int sy = g.sy, sz = g.sz;
    if (launch_interior(g, c.stream, [=] __device__(int i, int j, int k) {
            const int o = g.at(i, j, k);
#define U(di, dj, dk) u[o + (di) + sy * (dj) + sz * (dk)]
#define V(di, dj, dk) v[o + (di) + sy * (dj) + sz * (dk)]
#define W(di, dj, dk) w[o + (di) + sy * (dj) + sz * (dk)]
            const double keps = kt[o] / ep[o];
            const double k2eps = keps * kt[o];
            const double k2eps2 = keps * keps;
            const double k3eps2 = k2eps2 * kt[o];
            const double dudx = (U(1, 0, 0) - U(0, 0, 0)) / h;
            const double dudy = 0.25 * (U(0, 0, 0) - U(0, -1, 0) + U(1, 0, 0) - U(1, -1, 0) + U(0, 1, 0) - U(0, 0, 0) +
                                        U(1, 1, 0) - U(1, 0, 0)) /
                                h;
            const double dudz = 0.25 * (U(0, 0, 0) - U(0, 0, -1) + U(1, 0, 0) - U(1, 0, -1) + U(0, 0, 1) - U(0, 0, 0) +
                                        U(1, 0, 1) - U(1, 0, 0)) /
                                h;
            const double dvdx = 0.25 * (V(0, 0, 0) - V(-1, 0, 0) + V(1, 0, 0) - V(0, 0, 0) + V(0, 1, 0) - V(-1, 1, 0) +
                                        V(1, 1, 0) - V(0, 1, 0)) /
                                h;
            const double dvdy = (V(0, 1, 0) - V(0, 0, 0)) / h;
            const double dvdz = 0.25 * (V(0, 0, 0) - V(0, 0, -1) + V(0, 0, 1) - V(0, 0, 0) + V(0, 1, 0) - V(0, 1, -1) +
                                        V(0, 1, 1) - V(0, 1, 0)) /
                                h;
            const double dwdx = 0.25 * (W(0, 0, 0) - W(-1, 0, 0) + W(1, 0, 0) - W(0, 0, 0) + W(0, 0, 1) - W(-1, 0, 1) +
                                        W(1, 0, 1) - W(0, 0, 1)) /
                                h;
            const double dwdy = 0.25 * (W(0, 0, 0) - W(0, -1, 0) + W(0, 0, 1) - W(0, -1, 1) + W(0, 1, 0) - W(0, 0, 0) +
                                        W(0, 1, 1) - W(0, 0, 1)) /
                                h;
            const double dwdz = (W(0, 0, 1) - W(0, 0, 0)) / h;
#undef U
#undef V
#undef W
            const double ss11 = dudx, ss12 = 0.5 * (dudy + dvdx), ss13 = 0.5 * (dudz + dwdx), ss22 = dvdy,
                         ss23 = 0.5 * (dvdz + dwdy), ss33 = dwdz;
            const double sskk = ss11 + ss22 + ss33;
            const double sstar11 = ss11 - third * sskk, sstar22 = ss22 - third * sskk, sstar33 = ss33 - third * sskk;
            const double sstar = sqrt(sstar11 * sstar11 + sstar22 * sstar22 + sstar33 * sstar33 + 2.0 * ss12 * ss12 +
                                      2.0 * ss13 * ss13 + 2.0 * ss23 * ss23 + tiny);
            const double omeg12 = 0.5 * (dudy - dvdx), omeg13 = 0.5 * (dudz - dwdx), omeg23 = 0.5 * (dvdz - dwdy);
            const double omegastar = sqrt(2.0 * omeg12 * omeg12 + 2.0 * omeg13 * omeg13 + 2.0 * omeg23 * omeg23);
            const double wstar = (sstar11 * sstar11 * sstar11 + sstar22 * sstar22 * sstar22 + sstar33 * sstar33 * sstar33 +
                                  3.0 * sstar11 * ss12 * ss12 + 3.0 * ss12 * ss12 * sstar22 + 3.0 * sstar11 * ss13 * ss13 +
                                  6.0 * ss12 * ss13 * ss23 + 3.0 * ss13 * ss13 * sstar33 + 3.0 * sstar22 * ss23 * ss23 +
                                  3.0 * ss23 * ss23 * sstar33) /
                                 sstar / sstar / sstar;
            const double ustar = sqrt(sstar * sstar + omegastar * omegastar);
            double sixsqws = sixsq * wstar;
            if ((1.0 - fabs(sixsqws)) < 1.e-4) sixsqws = sixsqws < 0.0 ? -1.0 : 1.0;
            const double asstar = sixsq * cos(acos(sixsqws) / 3.0);
            const double cm = 1.0 / (6.5 + asstar * ustar * keps);
            cmu[o] = cm;
            mut[o] = cm * rh[o] * kt[o] * kt[o] / ep[o];
            double c2 = 1.0 - 9.0 * cm * cm * sstar * sstar * k2eps2;
            if (c2 < 0.0) c2 = 0.0;
            c2 = sqrt(c2) / (1.0 + 6.0 * sstar * k2eps2 * omegastar);
            u2p[o] = twothird * kt[o] - 2.0 * cm * k2eps * sstar11 + 4.0 * c2 * k3eps2 * (ss12 * omeg12 + ss13 * omeg13);
            v2p[o] = twothird * kt[o] - 2.0 * cm * k2eps * sstar22 - 4.0 * c2 * k3eps2 * (ss12 * omeg12 - ss23 * omeg23);
            w2p[o] = twothird * kt[o] - 2.0 * cm * k2eps * sstar33 - 4.0 * c2 * k3eps2 * (ss13 * omeg13 + ss23 * omeg23);
            uvp[o] = -(2.0 * cm * k2eps * ss12) +
                     2.0 * c2 * k3eps2 * (-(sstar11 * omeg12) + sstar22 * omeg12 + ss13 * omeg23 + ss23 * omeg13);
            uwp[o] = -(2.0 * cm * k2eps * ss13) +
                     2.0 * c2 * k3eps2 * (-(sstar11 * omeg13) + sstar33 * omeg13 - ss12 * omeg23 + ss23 * omeg12);
            vwp[o] = -(2.0 * cm * k2eps * ss23) +
                     2.0 * c2 * k3eps2 * (-(sstar22 * omeg23) + sstar33 * omeg23 - ss12 * omeg13 - ss13 * omeg12);
        }))
        return -1;
    for (double *a : {mut, u2p, v2p, w2p, uvp, uwp, vwp})
        if (op_bc_scalar(g, c.stream, a)) return -1;
    return 0;
}

// ---------------------------------------------------------------------------
// CFL (main.f:4848-4881) and checktime (main.f:2156-2175)

int op_cfl(Ctx &c, double *out_host)
{
    const Grid g = c.g;
    const double *u = c.u, *v = c.v, *w = c.w;
    // max is order independent: one pass over the three components
    if (reduce_interior<true>(g, c.stream, c.red_partial, c.dscal + 8, [=] __device__(int i, int j, int k) {
            const int o = g.at(i, j, k);
            return max_nan(max_nan(fabs(u[o]), fabs(v[o])), fabs(w[o]));
        }))
        return -1;
    DG_HIP(hipMemcpyAsync(c.hscal + 8, c.dscal + 8, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    DG_HIP(hipStreamSynchronize(c.stream));
    double m = c.hscal[8];
    if (m < 0.0) m = 0.0;  // CFL starts from 0 (main.f:4855)
    *out_host = m * c.s.tauh;
    // the reference carries a NaN velocity on silently (both tests of
    // checktime are false); here it ends the step with an error
    if (!std::isfinite(*out_host)) {
        set_error("CFL: non-finite velocity (max |u|,|v|,|w| = %g)", m);
        return -1;
    }
    return 0;
}

int op_checktime(Ctx &c, int t)
{
    (void)t;
    double cfl1;
    if (op_cfl(c, &cfl1)) return -1;
    c.cfl_last = cfl1;
    Scalars &s = c.s;
    if (cfl1 > s.maxcfl) {
        s.tau = s.tau * (s.maxcfl * DG_F4(0.8)) / cfl1;
        s.tauh = s.tau / s.h;
        s.tauh2 = (int)(s.tau / (s.h * s.h));
    }
    if (cfl1 <= s.maxcfl * 0.25) {
        cfl1 = s.maxcfl * 0.5;
        s.tau = s.tau * (s.maxcfl * 0.5) / cfl1;
        s.tauh = s.tau / s.h;
        s.tauh2 = (int)(s.tau / (s.h * s.h));
    }
    s.time = s.time + s.tau;
    return 0;
}

// ---------------------------------------------------------------------------
// CSF surface tension, csf (main.f:2363-2481), filterc27 (main.f:5473-5511)

// planes each thread of a column kernel walks
constexpr int COLUMN_KC = 8;

// f(P, o) on the interior for a stencil within the 3 x 3 x 3 block around
// the cell, with each thread walking COLUMN_KC planes up one column and
// keeping the 3 x 3 neighbourhoods of src in the planes below, at and
// above the cell in registers: P[dk + 1][(dj + 1) * 3 + di + 1] is
// src(i + di, j + dj, k + dk).  9 loads per cell instead of one per
// stencil point: the vector memory pipeline, not HBM, bounds the
// one-cell-per-thread form of such stencils.
template <class F>
__global__ void __launch_bounds__(256) k_column27(Grid g, const double *src, F f)
{
    const uint3 tile = xcd_tile();
    const int i = (int)(tile.x * blockDim.x + threadIdx.x) + 2;
    const int j = (int)(tile.y * blockDim.y + threadIdx.y) + 2;
    const int k0 = (int)tile.z * COLUMN_KC + 2;
    if (i > g.nx - 1 || j > g.ny - 1) return;
    const int k1 = k0 + COLUMN_KC - 1 < g.nz - 1 ? k0 + COLUMN_KC - 1 : g.nz - 1;
    const int sy = g.sy, sz = g.sz;
    double P[3][9];
    int o = g.at(i, j, k0);
    for (int q = 0; q < 9; q++) {
        const int d = (q % 3 - 1) + sy * (q / 3 - 1);
        P[0][q] = src[o - sz + d];
        P[1][q] = src[o + d];
    }
    for (int k = k0; k <= k1; k++, o += sz) {
        for (int q = 0; q < 9; q++) P[2][q] = src[o + sz + (q % 3 - 1) + sy * (q / 3 - 1)];
        f(P, o);
        for (int q = 0; q < 9; q++) {
            P[0][q] = P[1][q];
            P[1][q] = P[2][q];
        }
    }
}

template <class F>
static int launch_column27(const Grid &g, hipStream_t s, const double *src, F f)
{
    if (g.nx < 3 || g.ny < 3 || g.nz < 3) return 0;
    const dim3 grid((unsigned)((g.nx - 2 + BX - 1) / BX), (unsigned)((g.ny - 2 + BY - 1) / BY),
                    (unsigned)((g.nz - 2 + COLUMN_KC - 1) / COLUMN_KC));
    hipLaunchKernelGGL(k_column27<F>, grid, interior_block(), 0, s, g, src, f);
    DG_CHECK_LAUNCH();
    return 0;
}

#define DG_P27(di, dj, dk) P[(dk) + 1][((dj) + 1) * 3 + (di) + 1]

// the 27 values of a column kernel's block are one value, bit for bit
__device__ inline bool uniform27(const double (&P)[3][9])
{
    const long long v = __double_as_longlong(P[1][4]);
    bool u = true;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 9; q++) u = u && __double_as_longlong(P[p][q]) == v;
    return u;
}

int mg_filterc27(const Grid &g, hipStream_t s, const double *cc, double *cf)
{
    return launch_column27(g, s, cc, [=] __device__(const double (&P)[3][9], int o) {
        const double w0 = DG_F4(0.2729), w1 = DG_F4(0.08185), w2 = DG_F4(0.0178), w3 = DG_F4(.0028);
#define C DG_P27
        double t6 = C(-1, 0, 0) + C(1, 0, 0) + C(0, -1, 0) + C(0, 1, 0) + C(0, 0, 1) + C(0, 0, -1);
        double t12 = C(-1, -1, 0) + C(-1, 1, 0) + C(1, -1, 0) + C(1, 1, 0) + C(0, 1, 1) + C(0, -1, 1) + C(1, 0, 1) +
                     C(-1, 0, 1) + C(0, 1, -1) + C(0, -1, -1) + C(1, 0, -1) + C(-1, 0, -1);
        double t8 = C(1, 1, 1) + C(1, 1, -1) + C(1, -1, 1) + C(1, -1, -1) + C(-1, 1, 1) + C(-1, 1, -1) + C(-1, -1, 1) +
                    C(-1, -1, -1);
        double v = w0 * C(0, 0, 0) + w1 * t6 + w2 * t12 + w3 * t8;
#undef C
        if (v > 1.0) v = 1.0;
        cf[o] = v;
    });
}

// dst[q] = src[q] on the ghost shell of g (every cell with i, j or k on the
// boundary), for up to three array pairs; the edges are copied more than
// once, which a plain copy allows
static int copy_shell(const Grid &g, hipStream_t s, double *const dst[3], const double *const src[3], int n)
{
    double *d0 = dst[0], *d1 = n > 1 ? dst[1] : nullptr, *d2 = n > 2 ? dst[2] : nullptr;
    const double *s0 = src[0], *s1 = n > 1 ? src[1] : nullptr, *s2 = n > 2 ? src[2] : nullptr;
    auto cp = [=] __device__(int o) {
        d0[o] = s0[o];
        if (d1) d1[o] = s1[o];
        if (d2) d2[o] = s2[o];
    };
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (launch_face(ny, nz, s, [=] __device__(int j0, int k0) {
            cp(g.at(1, j0 + 1, k0 + 1));
            cp(g.at(nx, j0 + 1, k0 + 1));
        }))
        return -1;
    if (launch_face(nx, nz, s, [=] __device__(int i0, int k0) {
            cp(g.at(i0 + 1, 1, k0 + 1));
            cp(g.at(i0 + 1, ny, k0 + 1));
        }))
        return -1;
    return launch_face(nx, ny, s, [=] __device__(int i0, int j0) {
        cp(g.at(i0 + 1, j0 + 1, 1));
        cp(g.at(i0 + 1, j0 + 1, nz));
    });
}

// csf leaves w11, w22, w33 = s11, s22, s33 on every cell (main.f:2471-2473).
// With w_interior false only their ghost shells are written: the time step
// calls csf right before solvelpres, whose predictor rewrites the interior
// of all three before anything reads it.
int op_csf(Ctx &c, bool w_interior)
{
    const Grid g = c.g;
    Level &L = c.lev[c.ng - 1];
    double *ccc = L.cc, *w11 = L.u, *w22 = L.res, *w33 = L.rhs, *ccf = c.ccf;
    const size_t nb = sizeof(double) * g.n();
    hipStream_t s = c.stream;
    // ccf = filterc27(cc) through w22, then bc_out (main.f:2372-2379)
    if (c.s.sigma == 0.0) {
        if (mg_filterc27(g, s, ccc, w22)) return -1;
        DG_HIP(hipMemcpyAsync(ccf, w22, nb, hipMemcpyDeviceToDevice, s));
        if (op_bc_out(g, s, ccf)) return -1;
        double *arr[6] = {c.s11, c.s22, c.s33, c.s12, c.s13, c.s23};
        for (double *a : arr)
            if (mg_fill(g.n(), s, a, 0.0)) return -1;
        return 0;
    }
    // w22 is overwritten below, so the filter writes the interior of ccf
    // directly; its ghost shell takes w22's, which bc_out's edge copies read
    {
        double *dst[3] = {ccf, nullptr, nullptr};
        const double *src[3] = {w22, nullptr, nullptr};
        if (copy_shell(g, s, dst, src, 1)) return -1;
    }
    if (mg_filterc27(g, s, ccc, ccf)) return -1;
    if (op_bc_out(g, s, ccf)) return -1;
    const double sigh = 0.5 * (c.s.tau * c.s.tau) * c.s.sigma / (c.s.h * c.s.h * c.s.h);
    const int sy = g.sy, sz = g.sz;
    if (launch_column27(g, s, ccf, [=] __device__(const double (&P)[3][9], int o) {
            // a wavefront whose cells all see one value (bit for bit) on
            // their 3x3x3 blocks: every difference below is +0, every
            // normalised gradient +0 / sqrt(e50), and the result -(+0)
            if (__all(uniform27(P))) {
                w11[o] = -0.0;
                return;
            }
            const double e50 = 1.0e-50;
#define F DG_P27
            double nxu = sqrt(sq(F(1, 0, 0) - F(0, 0, 0)) + 0.0625 * sq(F(0, 1, 0) - F(0, -1, 0) + F(1, 1, 0) - F(1, -1, 0)) +
                              0.0625 * sq(F(0, 0, 1) - F(0, 0, -1) + F(1, 0, 1) - F(1, 0, -1)) + e50);
            double nxd = sqrt(sq(F(0, 0, 0) - F(-1, 0, 0)) + 0.0625 * sq(F(0, 1, 0) - F(0, -1, 0) + F(-1, 1, 0) - F(-1, -1, 0)) +
                              0.0625 * sq(F(0, 0, 1) - F(0, 0, -1) + F(-1, 0, 1) - F(-1, 0, -1)) + e50);
            double nyu = sqrt(sq(F(0, 1, 0) - F(0, 0, 0)) + 0.0625 * sq(F(1, 0, 0) - F(-1, 0, 0) + F(1, 1, 0) - F(-1, 1, 0)) +
                              0.0625 * sq(F(0, 0, 1) - F(0, 0, -1) + F(0, 1, 1) - F(0, 1, -1)) + e50);
            double nyd = sqrt(sq(F(0, 0, 0) - F(0, -1, 0)) + 0.0625 * sq(F(1, 0, 0) - F(-1, 0, 0) + F(1, -1, 0) - F(-1, -1, 0)) +
                              0.0625 * sq(F(0, 0, 1) - F(0, 0, -1) + F(0, -1, 1) - F(0, -1, -1)) + e50);
            double nzu = sqrt(sq(F(0, 0, 1) - F(0, 0, 0)) + 0.0625 * sq(F(1, 0, 0) - F(-1, 0, 0) + F(1, 0, 1) - F(-1, 0, 1)) +
                              0.0625 * sq(F(0, 1, 0) - F(0, -1, 0) + F(0, 1, 1) - F(0, -1, 1)) + e50);
            double nzd = sqrt(sq(F(0, 0, 0) - F(0, 0, -1)) + 0.0625 * sq(F(1, 0, 0) - F(-1, 0, 0) + F(1, 0, -1) - F(-1, 0, -1)) +
                              0.0625 * sq(F(0, 1, 0) - F(0, -1, 0) + F(0, 1, -1) - F(0, -1, -1)) + e50);
            double d1x = (F(1, 0, 0) - F(0, 0, 0)) / nxu;
            double d2x = (F(0, 0, 0) - F(-1, 0, 0)) / nxd;
            double d1y = (F(0, 1, 0) - F(0, 0, 0)) / nyu;
            double d2y = (F(0, 0, 0) - F(0, -1, 0)) / nyd;
            double d1z = (F(0, 0, 1) - F(0, 0, 0)) / nzu;
            double d2z = (F(0, 0, 0) - F(0, 0, -1)) / nzd;
#undef F
            w11[o] = -(d1x - d2x + d1y - d2y + d1z - d2z);
        }))
        return -1;
    if (op_bc_out(g, s, w11)) return -1;
    double *s11 = c.s11, *s22 = c.s22, *s33 = c.s33;
    // the copies w = s of the interior are made by the kernel that writes s
    // (each cell reads only its own w11 there)
    const bool wi = w_interior;
    if (launch_interior(g, s, [=] __device__(int i, int j, int k) {
            const int o = g.at(i, j, k);
            const double cf = ccf[o], wv = w11[o];
            const double v11 = sigh * (cf - ccf[o - 1]) * wv;
            const double v22 = sigh * (cf - ccf[o - sy]) * wv;
            const double v33 = sigh * (cf - ccf[o - sz]) * wv;
            s11[o] = v11;
            s22[o] = v22;
            s33[o] = v33;
            if (wi) {
                w11[o] = v11;
                w22[o] = v22;
                w33[o] = v33;
            }
        }))
        return -1;
    {
        double *dst[3] = {w11, w22, w33};
        const double *src[3] = {s11, s22, s33};
        if (copy_shell(g, s, dst, src, 3)) return -1;
    }
    return op_bc_tensor(c);
}

// ---------------------------------------------------------------------------
// Interface heights, CalcH3d (main.f:1312-1810).  Per interior cell: the
// Youngs normal of the 27-point stencil, the plane cut of the unit cube
// holding min(c, 1-c) classified by the sorted normal components (the
// reference's cases: two-dimensional, triangle, quadrilateral A, pentagon
// and hexagon by Newton iteration, quadrilateral B), the six face heights
// mirrored for c > 0.5, assigned to the x/y/z faces by which component is
// the smallest and the middle one (the y heights of the last branch follow
// niz's sign, as in main.f:1691-1719), swapped for c <= 0.5 and clipped to
// [0, 1].  The reference STOPs when no case matches (main.f:1582-1583) or
// the middle component is not found (main.f:1641-1644): *stop takes 1 or 2.

__device__ inline void hset(bool negative, double a, double b, double &lo, double &hi)
{
    if (negative) { lo = a; hi = b; }
    else { lo = b; hi = a; }
}

__device__ inline double cube(double a) { return a * a * a; }

int op_calch3d(Ctx &c)
{
    const Grid g = c.g;
    if (!c.hxmin) {
        set_error("CalcH3d: the interface-height arrays are not allocated (hswitch=0 and no SGIP)");
        return -1;
    }
    double *hxl = c.hxmin, *hxh = c.hxmax, *hyl = c.hymin, *hyh = c.hymax, *hzl = c.hzmin, *hzh = c.hzmax;
    double *stop = c.dscal + 12;
    if (mg_fill(1, c.stream, stop, 0.0)) return -1;
    if (launch_column27(g, c.stream, c.lev[c.ng - 1].cc, [=] __device__(const double (&P)[3][9], int o) {
#define C DG_P27
            double mm1 = C(-1, -1, -1) + C(-1, -1, 1) + C(-1, 1, -1) + C(-1, 1, 1) +
                         2.0 * (C(-1, -1, 0) + C(-1, 1, 0) + C(-1, 0, -1) + C(-1, 0, 1)) + 4.0 * C(-1, 0, 0);
            double mm2 = C(1, -1, -1) + C(1, -1, 1) + C(1, 1, -1) + C(1, 1, 1) +
                         2.0 * (C(1, -1, 0) + C(1, 1, 0) + C(1, 0, -1) + C(1, 0, 1)) + 4.0 * C(1, 0, 0);
            const double normx = mm1 - mm2;
            mm1 = C(-1, -1, -1) + C(-1, -1, 1) + C(1, -1, -1) + C(1, -1, 1) +
                  2.0 * (C(-1, -1, 0) + C(1, -1, 0) + C(0, -1, -1) + C(0, -1, 1)) + 4.0 * C(0, -1, 0);
            mm2 = C(-1, 1, -1) + C(-1, 1, 1) + C(1, 1, -1) + C(1, 1, 1) +
                  2.0 * (C(-1, 1, 0) + C(1, 1, 0) + C(0, 1, -1) + C(0, 1, 1)) + 4.0 * C(0, 1, 0);
            const double normy = mm1 - mm2;
            mm1 = C(-1, -1, -1) + C(-1, 1, -1) + C(1, -1, -1) + C(1, 1, -1) +
                  2.0 * (C(-1, 0, -1) + C(1, 0, -1) + C(0, -1, -1) + C(0, 1, -1)) + 4.0 * C(0, 0, -1);
            mm2 = C(-1, -1, 1) + C(-1, 1, 1) + C(1, -1, 1) + C(1, 1, 1) +
                  2.0 * (C(-1, 0, 1) + C(1, 0, 1) + C(0, -1, 1) + C(0, 1, 1)) + 4.0 * C(0, 0, 1);
            const double normz = mm1 - mm2;
            const double c0 = C(0, 0, 0);
#undef C
            const double nn = sqrt(normx * normx + normy * normy + normz * normz + 1.0e-50);
            const double nix = normx / nn, niy = normy / nn, niz = normz / nn;
            const double ccnew = c0 > 0.5 ? 1.0 - c0 : c0;
            const double absxn = fabs(nix), absyn = fabs(niy), abszn = fabs(niz);
            const double n1 = dmin1(dmin1(absxn, absyn), abszn);
            const double n3 = dmax1(dmax1(absxn, absyn), abszn);
            const double n2 = absxn + absyn + abszn - n1 - n3;
            const double clim1 = sqrt(2.0 * ccnew * n3 * n2);
            double h1a, h1b, h2a, h2b, h3a, h3b, diag, s1, s2, s3;
            if (n1 == 0.0 && n2 > 0.0 && ccnew <= clim1) {
                s2 = sqrt(2.0 * ccnew * n3 / n2);
                s3 = sqrt(2.0 * ccnew * n2 / n3);
                h1a = s2 * s3 / 2.0; h1b = s2 * s3 / 2.0;
                h2a = s3; h2b = 0.0; h3a = s2; h3b = 0.0;
            } else if (n1 == 0.0 && n2 > 0.0 && ccnew > clim1) {
                h1a = c0; h1b = c0;
                h2a = c0 + n2 / (2.0 * n3); h2b = c0 - n2 / (2.0 * n3);
                h3a = 1.0; h3b = 0.0;
            } else if (n1 == 0.0 && n2 == 0.0) {
                h1a = ccnew; h1b = ccnew; h2a = ccnew; h2b = ccnew; h3a = 1.0; h3b = 0.0;
            } else if (6.0 * n2 * n3 * ccnew < n1 * n1) {
                diag = pow(6.0 * n1 * n2 * n3 * ccnew, 1.0 / 3.0);
                s1 = diag / n1; s2 = diag / n2; s3 = diag / n3;
                h1a = (s2 * s3) / 2.0; h1b = 0.0; h2a = (s1 * s3) / 2.0; h2b = 0.0; h3a = (s1 * s2) / 2.0; h3b = 0.0;
            } else if (6.0 * n2 * n3 * ccnew < 3.0 * (n2 * n2) - 3.0 * n1 * n2 + n1 * n1) {
                diag = (3.0 * n1 + sqrt(72.0 * n2 * n3 * ccnew - 3.0 * (n1 * n1))) / 6.0;
                s1 = diag / n1; s2 = diag / n2; s3 = diag / n3;
                h1a = (s2 * s3) / 2.0;
                h1b = (s2 * s3) / 2.0 * sq((s1 - 1.0) / s1);
                h2a = s3 * (2.0 * s1 - 1.0) / (2.0 * s1); h2b = 0.0;
                h3a = s2 * (2.0 * s1 - 1.0) / (2.0 * s1); h3b = 0.0;
            } else if ((n1 + n2 >= n3 && 6.0 * n1 * n2 * n3 * ccnew < n3 * n3 * n3 - cube(n3 - n1) - cube(n3 - n2)) ||
                       (n1 + n2 < n3 && 2.0 * n3 * ccnew < n1 + n2)) {
                const double rsixn3 = 1.0 / (6.0 * n1 * n2 * n3);
                diag = n2;
                for (int loop = 1; loop <= 1000; loop++) {
                    const double delta = -(rsixn3 * (cube(diag) - cube(diag - n1) - cube(diag - n2)) - ccnew) /
                                         (3.0 * rsixn3 * (diag * diag - sq(diag - n1) - sq(diag - n2)));
                    diag = diag + delta;
                    if (fabs(delta) < 1.0e-6) break;
                }
                s1 = diag / n1; s2 = diag / n2; s3 = diag / n3;
                h1a = s3 * (2.0 * s2 - 1.0) / (2.0 * s2);
                h1b = (s2 * s3) / 2.0 * sq((s1 - 1.0) / s1);
                h2a = s3 * (2.0 * s1 - 1.0) / (2.0 * s1);
                h2b = (s1 * s3) / 2.0 * sq((s2 - 1.0) / s2);
                h3a = 2.0 * (s1 + s2) - s1 / s2 - s2 / s1 - s1 * s2 - 1.0; h3b = 0.0;
            } else if (n1 + n2 > n3 && 6.0 * n1 * n2 * n3 * ccnew >= n3 * n3 * n3 - cube(n3 - n1) - cube(n3 - n2)) {
                const double rsixn3 = 1.0 / (6.0 * n1 * n2 * n3);
                diag = n3;
                for (int loop = 1; loop <= 1000; loop++) {
                    const double delta =
                        -(rsixn3 * (cube(diag) - cube(diag - n1) - cube(diag - n2) - cube(diag - n3)) - ccnew) /
                        (3.0 * rsixn3 * (diag * diag - sq(diag - n1) - sq(diag - n2) - sq(diag - n3)));
                    diag = diag + delta;
                    if (fabs(delta) < 1.0e-6) break;
                }
                s1 = diag / n1; s2 = diag / n2; s3 = diag / n3;
                h1a = 2.0 * (s2 + s3) - s2 / s3 - s3 / s2 - s2 * s3 - 1.0;
                h1b = (s2 * s3) / 2.0 * sq((s1 - 1.0) / s1);
                h2a = 2.0 * (s1 + s3) - s1 / s3 - s3 / s1 - s1 * s3 - 1.0;
                h2b = (s1 * s3) / 2.0 * sq((s2 - 1.0) / s2);
                h3a = 2.0 * (s1 + s2) - s1 / s2 - s2 / s1 - s1 * s2 - 1.0;
                h3b = (s1 * s2) / 2.0 * sq((s3 - 1.0) / s3);
            } else if (n1 + n2 < n3 && 6.0 * n1 * n2 * n3 * ccnew >= cube(n1 + n2) - n1 * n1 * n1 - n2 * n2 * n2) {
                diag = 0.5 * (n1 + n2) + ccnew * n3;
                s1 = diag / n1; s2 = diag / n2; s3 = diag / n3;
                h1a = s3 * (2.0 * s2 - 1.0) / (2.0 * s2);
                h1b = s3 * (2.0 * s2 - 1.0) * (s1 - 1.0) / (2.0 * s1 * s2);
                h2a = s3 * (2.0 * s1 - 1.0) / (2.0 * s1);
                h2b = s3 * (2.0 * s1 - 1.0) * (s2 - 1.0) / (2.0 * s1 * s2);
                h3a = 1.0; h3b = 0.0;
            } else {
                *stop = 1.0;  // NO MATCH CASES FOUND
                return;
            }
            if (c0 > 0.5) {
                h1a = 1.0 - h1a; h1b = 1.0 - h1b; h2a = 1.0 - h2a;
                h2b = 1.0 - h2b; h3a = 1.0 - h3a; h3b = 1.0 - h3b;
            }
            double xl, xh, yl, yh, zl, zh;
            if (absxn == n1) {
                hset(nix < 0, h1a, h1b, xl, xh);
                if (fabs(absyn - n2) < 1.0e-6) {
                    hset(niy < 0, h2a, h2b, yl, yh);
                    hset(niz < 0, h3a, h3b, zl, zh);
                } else if (fabs(abszn - n2) < 1.0e-6) {
                    hset(niy < 0, h3a, h3b, yl, yh);
                    hset(niz < 0, h2a, h2b, zl, zh);
                } else {
                    *stop = 2.0;
                    return;
                }
            } else if (absyn == n1) {
                hset(niy < 0, h1a, h1b, yl, yh);
                if (fabs(absxn - n2) < 1.0e-6) {
                    hset(nix < 0, h2a, h2b, xl, xh);
                    hset(niz < 0, h3a, h3b, zl, zh);
                } else if (fabs(abszn - n2) < 1.0e-6) {
                    hset(nix < 0, h3a, h3b, xl, xh);
                    hset(niz < 0, h2a, h2b, zl, zh);
                } else {
                    *stop = 2.0;
                    return;
                }
            } else if (abszn == n1) {
                hset(niz < 0, h1a, h1b, zl, zh);
                if (fabs(absxn - n2) < 1.0e-6) {
                    hset(nix < 0, h2a, h2b, xl, xh);
                    hset(niz < 0, h3a, h3b, yl, yh);
                } else if (fabs(absyn - n2) < 1.0e-6) {
                    hset(nix < 0, h3a, h3b, xl, xh);
                    hset(niz < 0, h2a, h2b, yl, yh);
                } else {
                    *stop = 2.0;
                    return;
                }
            } else {
                // a NaN normal: the reference assigns nothing here
                xl = hxl[o]; xh = hxh[o]; yl = hyl[o]; yh = hyh[o]; zl = hzl[o]; zh = hzh[o];
            }
            if (c0 <= 0.5) {
                double t;
                t = xl; xl = xh; xh = t;
                t = yl; yl = yh; yh = t;
                t = zl; zl = zh; zh = t;
            }
            auto clip = [](double v) {
                if (v < 0.0) v = 0.0;
                if (v > 1.0) v = 1.0;
                return v;
            };
            hxl[o] = clip(xl); hxh[o] = clip(xh); hyl[o] = clip(yl);
            hyh[o] = clip(yh); hzl[o] = clip(zl); hzh[o] = clip(zh);
        }))
        return -1;
    DG_HIP(hipMemcpyAsync(c.hscal + 12, stop, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    DG_HIP(hipStreamSynchronize(c.stream));
    if (c.hscal[12] == 1.0) {
        set_error("STOP: NO MATCH CASES FOUND (CalcH3d, main.f:1582)");
        return -1;
    }
    if (c.hscal[12] != 0.0) {
        set_error("STOP: PROBLEM finding MIN and MAX unit vector components (CalcH3d, main.f:1641)");
        return -1;
    }
    return 0;
}

// SGIP / SGIP2 (main.f:2835-3166): the filter passes exactly as the
// reference (ccf = ccc; nf times: w22 = filterc27(ccf), ccf = w22 over the
// whole array, bc_out(ccf)), then the curvature into w11 with the
// tangential sums of the x-down, y-up and z-down faces in SGIP's own order
// (main.f:2878-2915; csf adds them the other way round), bc_out(w11).
static int sgip_front(Ctx &c, int nf)
{
    const Grid g = c.g;
    Level &L = c.lev[c.ng - 1];
    double *ccc = L.cc, *w11 = L.u, *w22 = L.res, *ccf = c.ccf;
    const size_t nb = sizeof(double) * g.n();
    hipStream_t s = c.stream;
    DG_HIP(hipMemcpyAsync(ccf, ccc, nb, hipMemcpyDeviceToDevice, s));
    for (int q = 0; q < nf; q++) {
        if (mg_filterc27(g, s, ccf, w22)) return -1;
        DG_HIP(hipMemcpyAsync(ccf, w22, nb, hipMemcpyDeviceToDevice, s));
        if (op_bc_out(g, s, ccf)) return -1;
    }
    if (c.s.sigma == 0.0) return 0;
    if (launch_column27(g, s, ccf, [=] __device__(const double (&P)[3][9], int o) {
            // a wavefront whose cells all see one value (bit for bit) on
            // their 3x3x3 blocks: every difference below is +0, every
            // normalised gradient +0 / sqrt(e50), and the result -(+0)
            if (__all(uniform27(P))) {
                w11[o] = -0.0;
                return;
            }
            const double e50 = 1.0e-50;
#define F DG_P27
            double nxu = sqrt(sq(F(1, 0, 0) - F(0, 0, 0)) + 0.0625 * sq(F(0, 1, 0) - F(0, -1, 0) + F(1, 1, 0) - F(1, -1, 0)) +
                              0.0625 * sq(F(0, 0, 1) - F(0, 0, -1) + F(1, 0, 1) - F(1, 0, -1)) + e50);
            double nxd = sqrt(sq(F(0, 0, 0) - F(-1, 0, 0)) + 0.0625 * sq(F(-1, 1, 0) - F(-1, -1, 0) + F(0, 1, 0) - F(0, -1, 0)) +
                              0.0625 * sq(F(-1, 0, 1) - F(-1, 0, -1) + F(0, 0, 1) - F(0, 0, -1)) + e50);
            double nyu = sqrt(sq(F(0, 1, 0) - F(0, 0, 0)) + 0.0625 * sq(F(1, 1, 0) - F(-1, 1, 0) + F(1, 0, 0) - F(-1, 0, 0)) +
                              0.0625 * sq(F(0, 1, 1) - F(0, 1, -1) + F(0, 0, 1) - F(0, 0, -1)) + e50);
            double nyd = sqrt(sq(F(0, 0, 0) - F(0, -1, 0)) + 0.0625 * sq(F(1, 0, 0) - F(-1, 0, 0) + F(1, -1, 0) - F(-1, -1, 0)) +
                              0.0625 * sq(F(0, 0, 1) - F(0, 0, -1) + F(0, -1, 1) - F(0, -1, -1)) + e50);
            double nzu = sqrt(sq(F(0, 0, 1) - F(0, 0, 0)) + 0.0625 * sq(F(1, 0, 0) - F(-1, 0, 0) + F(1, 0, 1) - F(-1, 0, 1)) +
                              0.0625 * sq(F(0, 1, 0) - F(0, -1, 0) + F(0, 1, 1) - F(0, -1, 1)) + e50);
            double nzd = sqrt(sq(F(0, 0, 0) - F(0, 0, -1)) + 0.0625 * sq(F(1, 0, -1) - F(-1, 0, -1) + F(1, 0, 0) - F(-1, 0, 0)) +
                              0.0625 * sq(F(0, 1, -1) - F(0, -1, -1) + F(0, 1, 0) - F(0, -1, 0)) + e50);
            double d1x = (F(1, 0, 0) - F(0, 0, 0)) / nxu;
            double d2x = (F(0, 0, 0) - F(-1, 0, 0)) / nxd;
            double d1y = (F(0, 1, 0) - F(0, 0, 0)) / nyu;
            double d2y = (F(0, 0, 0) - F(0, -1, 0)) / nyd;
            double d1z = (F(0, 0, 1) - F(0, 0, 0)) / nzu;
            double d2z = (F(0, 0, 0) - F(0, 0, -1)) / nzd;
#undef F
            w11[o] = -(d1x - d2x + d1y - d2y + d1z - d2z);
        }))
        return -1;
    return op_bc_out(g, s, w11);
}

// sigma = 0: the six stresses are zeroed and nothing is copied (the w's keep
// what they hold: w22 the last filter pass)
static int zero_stresses(Ctx &c)
{
    double *arr[6] = {c.s11, c.s22, c.s33, c.s12, c.s13, c.s23};
    for (double *a : arr)
        if (mg_fill(c.g.n(), c.stream, a, 0.0)) return -1;
    return 0;
}

// w11, w22, w33 = s11, s22, s33 over the whole array (main.f:3017-3019,
// 3156-3158), or only their ghost shells when the caller rewrites the
// interiors before reading them (the time step's predictor)
static int copy_w_from_s(Ctx &c, bool w_interior)
{
    Level &L = c.lev[c.ng - 1];
    double *dst[3] = {L.u, L.res, L.rhs};
    const double *src[3] = {c.s11, c.s22, c.s33};
    if (w_interior) {
        const size_t nb = sizeof(double) * c.g.n();
        for (int q = 0; q < 3; q++) DG_HIP(hipMemcpyAsync(dst[q], src[q], nb, hipMemcpyDeviceToDevice, c.stream));
        return 0;
    }
    return copy_shell(c.g, c.stream, dst, src, 3);
}

// SIND / COSD of the reference (the Fortran 2023 intrinsics its build maps
// dsind / dcosd to): sine and cosine of the argument times pi/180
__device__ inline double sind_deg(double x) { return sin(x * (3.14159265358979323846 / 180.0)); }
__device__ inline double cosd_deg(double x) { return cos(x * (3.14159265358979323846 / 180.0)); }

// SGIP (main.f:2835-3027): S_dd = sigh (w11 + w11_down)(ccf_up - ccf_down)
// |L - L_down| / 2 with L the mean CalcH3d height along d; on j = 2 the
// contact-angle form, theta from the contact-line velocity (115 or 45
// degrees beyond |v_n| > 0.1, else 350 v_n + 80); bc_out of S11..S33; no
// bc_tensor.
int op_sgip(Ctx &c, bool w_interior)
{
    if (!c.hxmin) {
        set_error("SGIP: the interface-height arrays are not allocated");
        return -1;
    }
    if (sgip_front(c, 1)) return -1;
    if (c.s.sigma == 0.0) return zero_stresses(c);
    const Grid g = c.g;
    Level &L = c.lev[c.ng - 1];
    const double *w11 = L.u, *ccf = c.ccf, *u = c.u, *w = c.w;
    const double *hxl = c.hxmin, *hxh = c.hxmax, *hyl = c.hymin, *hyh = c.hymax, *hzl = c.hzmin, *hzh = c.hzmax;
    double *s11 = c.s11, *s22 = c.s22, *s33 = c.s33;
    const double sigh = 0.5 * (c.s.tau * c.s.tau) * c.s.sigma / (c.s.h * c.s.h * c.s.h);
    const int sy = g.sy, sz = g.sz;
    if (launch_interior(g, c.stream, [=] __device__(int i, int j, int k) {
            const int o = g.at(i, j, k);
            const double Lx = (hxh[o] + hxl[o]) / 2.0, Lxd = (hxh[o - 1] + hxl[o - 1]) / 2.0;
            const double Ly = (hyh[o] + hyl[o]) / 2.0, Lyd = (hyh[o - sy] + hyl[o - sy]) / 2.0;
            const double Lz = (hzh[o] + hzl[o]) / 2.0, Lzd = (hzh[o - sz] + hzl[o - sz]) / 2.0;
            const double wv = w11[o];
            double v11 = sigh * (wv + w11[o - 1]) * (ccf[o + 1] - ccf[o - 1]) * fabs(Lx - Lxd) / 2.0;
            double v22 = sigh * (wv + w11[o - sy]) * (ccf[o + sy] - ccf[o - sy]) * fabs(Ly - Lyd) / 2.0;
            double v33 = sigh * (wv + w11[o - sz]) * (ccf[o + sz] - ccf[o - sz]) * fabs(Lz - Lzd) / 2.0;
            if (j == 2) {
                const double mx = ccf[o + 1] - ccf[o - 1], mz = ccf[o + sz] - ccf[o - sz];
                const double mmx = mx / sqrt(mz * mz + mx * mx + 1.0e-50);
                const double mmz = mz / sqrt(mz * mz + mx * mx + 1.0e-50);
                double normv = mmx * u[o] + mmz * w[o];
                normv = -normv;
                double theta = 0.0;
                if (fabs(normv) > DG_F4(0.1)) {
                    if (normv > 0) theta = 115.0;
                    if (normv < 0) theta = 45.0;
                } else {
                    theta = DG_F4(350.0) * normv + 80.0;
                }
                const double my = ccf[o + sy] - ccf[o - sy];
                const double mmy = cosd_deg(theta);
                v11 = sigh * (wv + w11[o - 1]) * mx * sind_deg(theta) / 2.0;
                v22 = sigh * (wv + w11[o - sy]) * mmy * sqrt(mz * mz + mx * mx + my * my) * -1.0;
                v33 = sigh * (wv + w11[o - sz]) * mz * sind_deg(theta) / 2.0;
            }
            s11[o] = v11;
            s22[o] = v22;
            s33[o] = v33;
        }))
        return -1;
    if (op_bc_out(g, c.stream, s11) || op_bc_out(g, c.stream, s22) || op_bc_out(g, c.stream, s33)) return -1;
    return copy_w_from_s(c, w_interior);
}

// SGIP2 (main.f:3029-3166): three filter passes, sigh = 0.5 tau sigma / h^2,
// S_dd = sigh (w11_up + w11_down)(ccf_up - ccf_down)(L_up - L_down) / 4; no
// bc_out of S, no bc_tensor.
int op_sgip2(Ctx &c, bool w_interior)
{
    if (!c.hxmin) {
        set_error("SGIP2: the interface-height arrays are not allocated");
        return -1;
    }
    if (sgip_front(c, 3)) return -1;
    if (c.s.sigma == 0.0) return zero_stresses(c);
    const Grid g = c.g;
    Level &L = c.lev[c.ng - 1];
    const double *w11 = L.u, *ccf = c.ccf;
    const double *hxl = c.hxmin, *hxh = c.hxmax, *hyl = c.hymin, *hyh = c.hymax, *hzl = c.hzmin, *hzh = c.hzmax;
    double *s11 = c.s11, *s22 = c.s22, *s33 = c.s33;
    const double sigh = 0.5 * c.s.tau * c.s.sigma / (c.s.h * c.s.h);
    const int sy = g.sy, sz = g.sz;
    if (launch_interior(g, c.stream, [=] __device__(int i, int j, int k) {
            const int o = g.at(i, j, k);
            const double Lxu = (hxh[o + 1] + hxl[o + 1]) / 2.0, Lxd = (hxh[o - 1] + hxl[o - 1]) / 2.0;
            const double Lyu = (hyh[o + sy] + hyl[o + sy]) / 2.0, Lyd = (hyh[o - sy] + hyl[o - sy]) / 2.0;
            const double Lzu = (hzh[o + sz] + hzl[o + sz]) / 2.0, Lzd = (hzh[o - sz] + hzl[o - sz]) / 2.0;
            s11[o] = sigh * (w11[o + 1] + w11[o - 1]) * (ccf[o + 1] - ccf[o - 1]) * (Lxu - Lxd) * 0.25;
            s22[o] = sigh * (w11[o + sy] + w11[o - sy]) * (ccf[o + sy] - ccf[o - sy]) * (Lyu - Lyd) * 0.25;
            s33[o] = sigh * (w11[o + sz] + w11[o - sz]) * (ccf[o + sz] - ccf[o - sz]) * (Lzu - Lzd) * 0.25;
        }))
        return -1;
    return copy_w_from_s(c, w_interior);
}

// ---------------------------------------------------------------------------
// solvelpres (main.f:3236-3439) with convective (main.f:3441-3580).
// ftr = 0 (main.f:3255) and turbst = 0 on this path, so the mut and turbst
// terms add +0.0 and are not evaluated; the Reynolds-stress arrays of
// convective are zero unless iturb = 4.

static int stress(Ctx &c)
{
    const Grid g = c.g;
    const double *cc = c.lev[c.ng - 1].cc, *u = c.u, *v = c.v, *w = c.w;
    double *s11 = c.s11, *s22 = c.s22, *s33 = c.s33, *s12 = c.s12, *s13 = c.s13, *s23 = c.s23;
    const double tauh2 = (double)c.s.tauh2;
    const double r20 = tauh2 * c.s.mu2;
    const double r10 = 0.25 * tauh2 * (c.s.mu1 - c.s.mu2);
    const double r1 = 2.0 * tauh2 * (c.s.mu1 - c.s.mu2);
    const double r2 = 2.0 * tauh2 * c.s.mu2;
    const int sy = g.sy, sz = g.sz;
    // every value is computed before the first store: the compiler cannot
    // tell the stores from the input arrays, and a store between two loads
    // would make it reload and serialise them.  (Walking columns with the
    // plane below in registers fetched 15% less but took longer: fewer
    // waves in flight for a kernel that streams 16 arrays.)
    return launch_interior(g, c.stream, [=] __device__(int i, int j, int k) {
        const int o = g.at(i, j, k);
        double temp = r2 + r1 * cc[o];
        const double n11 = -s11[o] + temp * (u[o + 1] - u[o]);
        const double n22 = -s22[o] + temp * (v[o + sy] - v[o]);
        const double n33 = -s33[o] + temp * (w[o + sz] - w[o]);
        temp = r20 + r10 * (cc[o] + cc[o - 1] + cc[o - sy] + cc[o - 1 - sy]);
        const double n12 = -s12[o] + (temp) * (u[o] - u[o - sy] + v[o] - v[o - 1]);
        temp = r20 + r10 * (cc[o] + cc[o - 1] + cc[o - sz] + cc[o - 1 - sz]);
        const double n13 = -s13[o] + (temp) * (u[o] - u[o - sz] + w[o] - w[o - 1]);
        temp = r20 + r10 * (cc[o] + cc[o - sy] + cc[o - sz] + cc[o - sy - sz]);
        const double n23 = -s23[o] + (temp) * (v[o] - v[o - sz] + w[o] - w[o - sy]);
        s11[o] = n11;
        s22[o] = n22;
        s33[o] = n33;
        s12[o] = n12;
        s13[o] = n13;
        s23[o] = n23;
    });
}

int mg_abcei(const Grid &g, hipStream_t s, double *a, double *b, double *c, double *ei, const double *cc, double r1,
             double r2, void *split, int split_ct)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz, sy = g.sy, sz = g.sz;
    const int mh = (nx - 2) / 2, my = ny - 2;
    const size_t half = (size_t)mh * my * (nz - 2), n2 = 2 * half;
    // a, b, c and ei in one pass.  ei needs the coefficients of the faces
    // below the cell too; the thread forms them again with the same
    // expression, and at i = 2 takes bc_coef's a(1,j,k) = a(2,j,k).  The
    // other two ghost coefficients it reads, b(i,1,k) and c(i,j,1), have the
    // interior expression (main.f:4465-4494), so ei is what the reference's
    // separate loops give, bit for bit.
    if (launch_interior(g, s, [=] __device__(int i, int j, int k) {
            const int o = g.at(i, j, k);
            const double na = 1.0 / (r1 * (cc[o] + cc[o + 1]) + r2);
            const double nb = 1.0 / (r1 * (cc[o] + cc[o + sy]) + r2);
            const double nc = 1.0 / (r1 * (cc[o] + cc[o + sz]) + r2);
            const double aw = i > 2 ? 1.0 / (r1 * (cc[o - 1] + cc[o]) + r2) : na;
            const double bs = 1.0 / (r1 * (cc[o - sy] + cc[o]) + r2);
            const double cb = 1.0 / (r1 * (cc[o - sz] + cc[o]) + r2);
            a[o] = na;
            b[o] = nb;
            c[o] = nc;
            ei[o] = 1.0 / (na + aw + nb + bs + nc + cb);
            if (split) {  // the conversion of the split V-cycle's set-up (k_rb_coef)
                const size_t e = (size_t)((i + j + k) & 1) * half + ((size_t)(k - 2) * my + (j - 2)) * mh + ((i - 2) >> 1);
                if (split_ct == RB_CT_F16) {
                    _Float16 *p = (_Float16 *)split;
                    p[e] = (_Float16)na;
                    p[n2 + e] = (_Float16)nb;
                    p[2 * n2 + e] = (_Float16)nc;
                } else if (split_ct == RB_CT_F64) {
                    double *p = (double *)split;
                    p[e] = na;
                    p[n2 + e] = nb;
                    p[2 * n2 + e] = nc;
                } else {
                    float *p = (float *)split;
                    p[e] = (float)na;
                    p[n2 + e] = (float)nb;
                    p[2 * n2 + e] = (float)nc;
                }
            }
        }))
        return -1;
    // bc_coef (main.f:4465-4494): the three fills write different arrays
    if (launch_face(ny, nz, s, [=] __device__(int j0, int k0) {
            int j = j0 + 1, k = k0 + 1;
            a[g.at(1, j, k)] = a[g.at(2, j, k)];
            a[g.at(nx, j, k)] = a[g.at(nx - 1, j, k)];
        }))
        return -1;
    if (launch_face(nx - 2, nz - 2, s, [=] __device__(int i0, int k0) {
            int i = i0 + 2, k = k0 + 2;
            b[g.at(i, 1, k)] = 1.0 / (r1 * (cc[g.at(i, 1, k)] + cc[g.at(i, 2, k)]) + r2);
        }))
        return -1;
    return launch_face(nx - 2, ny, s, [=] __device__(int i0, int j0) {
        int i = i0 + 2, j = j0 + 1;
        c[g.at(i, j, 1)] = 1.0 / (r1 * (cc[g.at(i, j, 1)] + cc[g.at(i, j, 2)]) + r2);
    });
}

static int predictor(Ctx &c, double *w1, double *w2, double *w3)
{
    const Grid g = c.g;
    Level &L = c.lev[c.ng - 1];
    const double *a = L.a, *b = L.b, *cf = L.c, *u = c.u, *v = c.v, *w = c.w;
    const double *s11 = c.s11, *s22 = c.s22, *s33 = c.s33, *s12 = c.s12, *s13 = c.s13, *s23 = c.s23;
    const double tau = c.s.tau, h = c.s.h;
    const double stdgrav = c.s.g * (tau * tau) / h;
    const double horgrav = c.s.g1 * (tau * tau) / h;
    const int sy = g.sy, sz = g.sz;
    return launch_interior(g, c.stream, [=] __device__(int i, int j, int k) {
        const int o = g.at(i, j, k);
        const double n1 = u[o] + a[o - 1] * (s11[o] - s11[o - 1] + s12[o + sy] - s12[o] + s13[o + sz] - s13[o]) + horgrav;
        const double n2 = v[o] + b[o - sy] * (s12[o + 1] - s12[o] + s22[o] - s22[o - sy] + s23[o + sz] - s23[o]) + stdgrav;
        const double n3 = w[o] + cf[o - sz] * (s13[o + 1] - s13[o] + s23[o + sy] - s23[o] + s33[o] - s33[o - sz]);
        w1[o] = n1;
        w2[o] = n2;
        w3[o] = n3;
    });
}

// convective's stencil points of w1, w2, w3 within each plane, in window
// order: w1 (0,0) (-1,0) (1,0) (0,-1) (0,1) (1,-1); w2 (0,0) (-1,0) (1,0)
// (0,-1) (0,1) (-1,1); w3 (0,0) (-1,0) (1,0) (0,-1) (0,1)
constexpr int CV_DI[6] = {0, -1, 1, 0, 0, 1}, CV_DJ1[6] = {0, 0, 0, -1, 1, -1}, CV_DI2[6] = {0, -1, 1, 0, 0, -1},
              CV_DJ2[6] = {0, 0, 0, -1, 1, 1};
__host__ __device__ constexpr int cv_i1(int di, int dj)
{
    return dj == 0 ? (di == 0 ? 0 : di < 0 ? 1 : 2) : dj < 0 ? (di == 0 ? 3 : 5) : 4;
}
__host__ __device__ constexpr int cv_i2(int di, int dj)
{
    return dj == 0 ? (di == 0 ? 0 : di < 0 ? 1 : 2) : dj < 0 ? 3 : (di == 0 ? 4 : 5);
}
__host__ __device__ constexpr int cv_i3(int di, int dj) { return cv_i1(di, dj); }
static_assert(cv_i1(1, -1) == 5 && cv_i1(0, 1) == 4 && cv_i2(-1, 1) == 5 && cv_i2(0, -1) == 3, "window order");

// the three planes k-1, k, k+1 of the window around the cell
struct CvWin {
    double p1[3][6], p2[3][6], p3[3][5];
};

// f(W, o) on the interior with each thread walking COLUMN_KC planes up one
// column: the planes below and at the cell stay in registers from the
// previous step, so each plane of w1, w2, w3 is fetched once per column
// instead of once per plane that reads it (the one-cell-per-thread form
// fetched each about three times: a plane's tiles in flight on an XCD need
// three planes of three arrays, more than its L2)
template <class F>
__global__ void __launch_bounds__(256) k_convective_col(Grid g, const double *w1, const double *w2,
                                                        const double *w3, F f)
{
    const uint3 tile = xcd_tile();
    const int i = (int)(tile.x * blockDim.x + threadIdx.x) + 2;
    const int j = (int)(tile.y * blockDim.y + threadIdx.y) + 2;
    const int k0 = (int)tile.z * COLUMN_KC + 2;
    if (i > g.nx - 1 || j > g.ny - 1) return;
    const int k1 = k0 + COLUMN_KC - 1 < g.nz - 1 ? k0 + COLUMN_KC - 1 : g.nz - 1;
    const int sy = g.sy, sz = g.sz;
    CvWin W;
    auto load = [&](int slot, int oo) {
#pragma unroll
        for (int q = 0; q < 6; q++) {
            W.p1[slot][q] = w1[oo + CV_DI[q] + sy * CV_DJ1[q]];
            W.p2[slot][q] = w2[oo + CV_DI2[q] + sy * CV_DJ2[q]];
        }
#pragma unroll
        for (int q = 0; q < 5; q++) W.p3[slot][q] = w3[oo + CV_DI[q] + sy * CV_DJ1[q]];
    };
    int o = g.at(i, j, k0);
    load(0, o - sz);
    load(1, o);
    for (int k = k0; k <= k1; k++, o += sz) {
        load(2, o + sz);
        f(W, o);
#pragma unroll
        for (int q = 0; q < 6; q++) {
            W.p1[0][q] = W.p1[1][q]; W.p1[1][q] = W.p1[2][q];
            W.p2[0][q] = W.p2[1][q]; W.p2[1][q] = W.p2[2][q];
        }
#pragma unroll
        for (int q = 0; q < 5; q++) {
            W.p3[0][q] = W.p3[1][q]; W.p3[1][q] = W.p3[2][q];
        }
    }
}

template <class F>
static int launch_convective_col(const Grid &g, hipStream_t s, const double *w1, const double *w2, const double *w3,
                                 F f)
{
    if (g.nx < 3 || g.ny < 3 || g.nz < 3) return 0;
    const dim3 grid((unsigned)((g.nx - 2 + BX - 1) / BX), (unsigned)((g.ny - 2 + BY - 1) / BY),
                    (unsigned)((g.nz - 2 + COLUMN_KC - 1) / COLUMN_KC));
    hipLaunchKernelGGL(k_convective_col<F>, grid, interior_block(), 0, s, g, w1, w2, w3, f);
    DG_CHECK_LAUNCH();
    return 0;
}

// the 16 Reynolds-stress terms of one component, which the reference
// writes with alternating signs (+t0 - t1 + t2 ...): added one at a time to
// acc in that order
__device__ inline double add_alternating(double acc, const double (&t)[16], int first)
{
    for (int q = first; q < 16; q++) acc = (q & 1) ? acc - t[q] : acc + t[q];
    return acc;
}

// convective with the Reynolds-stress terms of the anisotropic model
// (iturb = 4, main.f:3469-3472, 3492-3495, 3513-3516 and 3536-3563), in the
// reference's order of evaluation: in the upwind form the terms are a
// parenthesised group of their own, in the central form they continue the
// 0.25 * (...) sum; one cell per thread
static int convective_reynolds(Ctx &c, const double *w1, const double *w2, const double *w3)
{
    const Grid g = c.g;
    double *u = c.u, *v = c.v, *w = c.w;
    const double tauh = c.s.tauh;
    const double *u2p = c.u2p, *v2p = c.v2p, *w2p = c.w2p, *uvp = c.uvp, *uwp = c.uwp, *vwp = c.vwp;
    const int sy = g.sy, sz = g.sz;
    const bool upwind = c.s.convectype == 1;
    return launch_interior(g, c.stream, [=] __device__(int i, int j, int k) {
        const int o = g.at(i, j, k);
#define A(a, di, dj, dk) a[o + (di) + sy * (dj) + sz * (dk)]
        const double ru[16] = {A(uvp, 0, 0, 0),  A(uvp, 0, 1, 0),   A(uvp, 0, -1, 0),  A(uvp, 0, 0, 0),
                               A(uvp, -1, 0, 0), A(uvp, -1, 1, 0),  A(uvp, -1, -1, 0), A(uvp, -1, 0, 0),
                               A(uwp, 0, 0, 0),  A(uwp, 0, 0, 1),   A(uwp, 0, 0, -1),  A(uwp, 0, 0, 0),
                               A(uwp, -1, 0, 0), A(uwp, -1, 0, 1),  A(uwp, -1, 0, -1), A(uwp, -1, 0, 0)};
        const double rv[16] = {A(uvp, -1, 0, 0),  A(uvp, 0, 0, 0),   A(uvp, 0, 0, 0),  A(uvp, 1, 0, 0),
                               A(uvp, -1, -1, 0), A(uvp, 0, -1, 0),  A(uvp, 0, -1, 0), A(uvp, 1, -1, 0),
                               A(vwp, 0, 0, -1),  A(vwp, 0, 0, 0),   A(vwp, 0, 0, 0),  A(vwp, 0, 0, 1),
                               A(vwp, 0, -1, -1), A(vwp, 0, -1, 0),  A(vwp, 0, -1, 0), A(vwp, 0, -1, 1)};
        const double rw[16] = {A(uwp, -1, 0, 0),  A(uwp, 0, 0, 0),   A(uwp, 0, 0, 0),  A(uwp, 1, 0, 0),
                               A(uwp, -1, 0, -1), A(uwp, 0, 0, -1),  A(uwp, 0, 0, -1), A(uwp, 1, 0, -1),
                               A(vwp, 0, -1, 0),  A(vwp, 0, 0, 0),   A(vwp, 0, 0, 0),  A(vwp, 0, 1, 0),
                               A(vwp, 0, -1, -1), A(vwp, 0, 0, -1),  A(vwp, 0, 0, -1), A(vwp, 0, 1, -1)};
        double nu, nv, nw;
        if (upwind) {
            double t1, t2, s1, s2, q1, q2;
            t1 = .5 * (A(w2, 0, 0, 0) + A(w2, -1, 0, 0));
            t2 = .5 * (A(w2, 0, 1, 0) + A(w2, -1, 1, 0));
            s1 = .5 * (A(w1, -1, 0, 0) + A(w1, 0, 0, 0));
            s2 = .5 * (A(w1, 0, 0, 0) + A(w1, 1, 0, 0));
            q1 = .5 * (A(w3, 0, 0, 0) + A(w3, -1, 0, 0));
            q2 = .5 * (A(w3, 0, 0, 1) + A(w3, -1, 0, 1));
            nu = A(w1, 0, 0, 0) +
                 tauh * (+dmax1(t1, 0.0) * A(w1, 0, -1, 0) + dmin1(t1, 0.0) * A(w1, 0, 0, 0) -
                         dmax1(t2, 0.0) * A(w1, 0, 0, 0) - dmin1(t2, 0.0) * A(w1, 0, 1, 0) +
                         dmax1(s1, 0.0) * A(w1, -1, 0, 0) + dmin1(s1, 0.0) * A(w1, 0, 0, 0) -
                         dmax1(s2, 0.0) * A(w1, 0, 0, 0) - dmin1(s2, 0.0) * A(w1, 1, 0, 0) +
                         dmax1(q1, 0.0) * A(w1, 0, 0, -1) + dmin1(q1, 0.0) * A(w1, 0, 0, 0) -
                         dmax1(q2, 0.0) * A(w1, 0, 0, 0) - dmin1(q2, 0.0) * A(w1, 0, 0, 1) + A(u2p, -1, 0, 0) -
                         A(u2p, 0, 0, 0) + 0.25 * add_alternating(ru[0], ru, 1));
            t1 = .5 * (A(w1, 0, 0, 0) + A(w1, 0, -1, 0));
            t2 = .5 * (A(w1, 1, 0, 0) + A(w1, 1, -1, 0));
            s1 = .5 * (A(w2, 0, -1, 0) + A(w2, 0, 0, 0));
            s2 = .5 * (A(w2, 0, 0, 0) + A(w2, 0, 1, 0));
            q1 = .5 * (A(w3, 0, 0, 0) + A(w3, 0, -1, 0));
            q2 = .5 * (A(w3, 0, 0, 1) + A(w3, 0, -1, 1));
            nv = A(w2, 0, 0, 0) +
                 tauh * (+dmax1(t1, 0.0) * A(w2, -1, 0, 0) + dmin1(t1, 0.0) * A(w2, 0, 0, 0) -
                         dmax1(t2, 0.0) * A(w2, 0, 0, 0) - dmin1(t2, 0.0) * A(w2, 1, 0, 0) +
                         dmax1(s1, 0.0) * A(w2, 0, -1, 0) + dmin1(s1, 0.0) * A(w2, 0, 0, 0) -
                         dmax1(s2, 0.0) * A(w2, 0, 0, 0) - dmin1(s2, 0.0) * A(w2, 0, 1, 0) +
                         dmax1(q1, 0.0) * A(w2, 0, 0, -1) + dmin1(q1, 0.0) * A(w2, 0, 0, 0) -
                         dmax1(q2, 0.0) * A(w2, 0, 0, 0) - dmin1(q2, 0.0) * A(w2, 0, 0, 1) + A(v2p, 0, -1, 0) -
                         A(v2p, 0, 0, 0) + 0.25 * add_alternating(rv[0], rv, 1));
            t1 = .5 * (A(w1, 0, 0, 0) + A(w1, 0, 0, -1));
            t2 = .5 * (A(w1, 1, 0, 0) + A(w1, 1, 0, -1));
            s1 = .5 * (A(w3, 0, 0, -1) + A(w3, 0, 0, 0));
            s2 = .5 * (A(w3, 0, 0, 0) + A(w3, 0, 0, 1));
            q1 = .5 * (A(w2, 0, 0, 0) + A(w2, 0, 0, -1));
            q2 = .5 * (A(w2, 0, 1, 0) + A(w2, 0, 1, -1));
            nw = A(w3, 0, 0, 0) +
                 tauh * (+dmax1(t1, 0.0) * A(w3, -1, 0, 0) + dmin1(t1, 0.0) * A(w3, 0, 0, 0) -
                         dmax1(t2, 0.0) * A(w3, 0, 0, 0) - dmin1(t2, 0.0) * A(w3, 1, 0, 0) +
                         dmax1(s1, 0.0) * A(w3, 0, 0, -1) + dmin1(s1, 0.0) * A(w3, 0, 0, 0) -
                         dmax1(s2, 0.0) * A(w3, 0, 0, 0) - dmin1(s2, 0.0) * A(w3, 0, 0, 1) +
                         dmax1(q1, 0.0) * A(w3, 0, -1, 0) + dmin1(q1, 0.0) * A(w3, 0, 0, 0) -
                         dmax1(q2, 0.0) * A(w3, 0, 0, 0) - dmin1(q2, 0.0) * A(w3, 0, 1, 0) + A(w2p, 0, 0, -1) -
                         A(w2p, 0, 0, 0) + 0.25 * add_alternating(rw[0], rw, 1));
        } else {
            nu = A(w1, 0, 0, 0) +
                 tauh * (0.25 * add_alternating(
                                    sq(A(w1, 0, 0, 0) + A(w1, -1, 0, 0)) - sq(A(w1, 1, 0, 0) + A(w1, 0, 0, 0)) +
                                        (A(w1, 0, -1, 0) + A(w1, 0, 0, 0)) * (A(w2, 0, 0, 0) + A(w2, -1, 0, 0)) -
                                        (A(w1, 0, 1, 0) + A(w1, 0, 0, 0)) * (A(w2, 0, 1, 0) + A(w2, -1, 1, 0)) +
                                        (A(w1, 0, 0, -1) + A(w1, 0, 0, 0)) * (A(w3, 0, 0, 0) + A(w3, -1, 0, 0)) -
                                        (A(w1, 0, 0, 1) + A(w1, 0, 0, 0)) * (A(w3, 0, 0, 1) + A(w3, -1, 0, 1)),
                                    ru, 0) +
                         A(u2p, -1, 0, 0) - A(u2p, 0, 0, 0));
            nv = A(w2, 0, 0, 0) +
                 tauh * (0.25 * add_alternating(
                                    sq(A(w2, 0, 0, 0) + A(w2, 0, -1, 0)) - sq(A(w2, 0, 1, 0) + A(w2, 0, 0, 0)) +
                                        (A(w2, -1, 0, 0) + A(w2, 0, 0, 0)) * (A(w1, 0, 0, 0) + A(w1, 0, -1, 0)) -
                                        (A(w2, 1, 0, 0) + A(w2, 0, 0, 0)) * (A(w1, 1, 0, 0) + A(w1, 1, -1, 0)) +
                                        (A(w2, 0, 0, -1) + A(w2, 0, 0, 0)) * (A(w3, 0, 0, 0) + A(w3, 0, -1, 0)) -
                                        (A(w2, 0, 0, 1) + A(w2, 0, 0, 0)) * (A(w3, 0, 0, 1) + A(w3, 0, -1, 1)),
                                    rv, 0) +
                         A(v2p, 0, -1, 0) - A(v2p, 0, 0, 0));
            nw = A(w3, 0, 0, 0) +
                 tauh * (0.25 * add_alternating(
                                    sq(A(w3, 0, 0, 0) + A(w3, 0, 0, -1)) - sq(A(w3, 0, 0, 1) + A(w3, 0, 0, 0)) +
                                        (A(w3, -1, 0, 0) + A(w3, 0, 0, 0)) * (A(w1, 0, 0, 0) + A(w1, 0, 0, -1)) -
                                        (A(w3, 1, 0, 0) + A(w3, 0, 0, 0)) * (A(w1, 1, 0, 0) + A(w1, 1, 0, -1)) +
                                        (A(w3, 0, -1, 0) + A(w3, 0, 0, 0)) * (A(w2, 0, 0, 0) + A(w2, 0, 0, -1)) -
                                        (A(w3, 0, 1, 0) + A(w3, 0, 0, 0)) * (A(w2, 0, 1, 0) + A(w2, 0, 1, -1)),
                                    rw, 0) +
                         A(w2p, 0, 0, -1) - A(w2p, 0, 0, 0));
        }
#undef A
        u[o] = nu;
        v[o] = nv;
        w[o] = nw;
    });
}

static int convective(Ctx &c, const double *w1, const double *w2, const double *w3)
{
    const Grid g = c.g;
    double *u = c.u, *v = c.v, *w = c.w;
    const double tauh = c.s.tauh;
    if (c.s.iturb == 4 && c.u2p && (c.s.convectype == 1 || c.s.convectype == 2))
        return convective_reynolds(c, w1, w2, w3);
    if (c.s.convectype == 1) {
        return launch_convective_col(g, c.stream, w1, w2, w3, [=] __device__(const CvWin &W, int o) {
#define A1(di, dj, dk) W.p1[(dk) + 1][cv_i1(di, dj)]
#define A2(di, dj, dk) W.p2[(dk) + 1][cv_i2(di, dj)]
#define A3(di, dj, dk) W.p3[(dk) + 1][cv_i3(di, dj)]
            double t1, t2, s1, s2, q1, q2;
            t1 = .5 * (A2(0, 0, 0) + A2(-1, 0, 0));
            t2 = .5 * (A2(0, 1, 0) + A2(-1, 1, 0));
            s1 = .5 * (A1(-1, 0, 0) + A1(0, 0, 0));
            s2 = .5 * (A1(0, 0, 0) + A1(1, 0, 0));
            q1 = .5 * (A3(0, 0, 0) + A3(-1, 0, 0));
            q2 = .5 * (A3(0, 0, 1) + A3(-1, 0, 1));
            const double nu = A1(0, 0, 0) + tauh * (+dmax1(t1, 0.0) * A1(0, -1, 0) + dmin1(t1, 0.0) * A1(0, 0, 0) -
                                         dmax1(t2, 0.0) * A1(0, 0, 0) - dmin1(t2, 0.0) * A1(0, 1, 0) +
                                         dmax1(s1, 0.0) * A1(-1, 0, 0) + dmin1(s1, 0.0) * A1(0, 0, 0) -
                                         dmax1(s2, 0.0) * A1(0, 0, 0) - dmin1(s2, 0.0) * A1(1, 0, 0) +
                                         dmax1(q1, 0.0) * A1(0, 0, -1) + dmin1(q1, 0.0) * A1(0, 0, 0) -
                                         dmax1(q2, 0.0) * A1(0, 0, 0) - dmin1(q2, 0.0) * A1(0, 0, 1));
            t1 = .5 * (A1(0, 0, 0) + A1(0, -1, 0));
            t2 = .5 * (A1(1, 0, 0) + A1(1, -1, 0));
            s1 = .5 * (A2(0, -1, 0) + A2(0, 0, 0));
            s2 = .5 * (A2(0, 0, 0) + A2(0, 1, 0));
            q1 = .5 * (A3(0, 0, 0) + A3(0, -1, 0));
            q2 = .5 * (A3(0, 0, 1) + A3(0, -1, 1));
            const double nv = A2(0, 0, 0) + tauh * (+dmax1(t1, 0.0) * A2(-1, 0, 0) + dmin1(t1, 0.0) * A2(0, 0, 0) -
                                         dmax1(t2, 0.0) * A2(0, 0, 0) - dmin1(t2, 0.0) * A2(1, 0, 0) +
                                         dmax1(s1, 0.0) * A2(0, -1, 0) + dmin1(s1, 0.0) * A2(0, 0, 0) -
                                         dmax1(s2, 0.0) * A2(0, 0, 0) - dmin1(s2, 0.0) * A2(0, 1, 0) +
                                         dmax1(q1, 0.0) * A2(0, 0, -1) + dmin1(q1, 0.0) * A2(0, 0, 0) -
                                         dmax1(q2, 0.0) * A2(0, 0, 0) - dmin1(q2, 0.0) * A2(0, 0, 1));
            t1 = .5 * (A1(0, 0, 0) + A1(0, 0, -1));
            t2 = .5 * (A1(1, 0, 0) + A1(1, 0, -1));
            s1 = .5 * (A3(0, 0, -1) + A3(0, 0, 0));
            s2 = .5 * (A3(0, 0, 0) + A3(0, 0, 1));
            q1 = .5 * (A2(0, 0, 0) + A2(0, 0, -1));
            q2 = .5 * (A2(0, 1, 0) + A2(0, 1, -1));
            const double nw = A3(0, 0, 0) + tauh * (+dmax1(t1, 0.0) * A3(-1, 0, 0) + dmin1(t1, 0.0) * A3(0, 0, 0) -
                                         dmax1(t2, 0.0) * A3(0, 0, 0) - dmin1(t2, 0.0) * A3(1, 0, 0) +
                                         dmax1(s1, 0.0) * A3(0, 0, -1) + dmin1(s1, 0.0) * A3(0, 0, 0) -
                                         dmax1(s2, 0.0) * A3(0, 0, 0) - dmin1(s2, 0.0) * A3(0, 0, 1) +
                                         dmax1(q1, 0.0) * A3(0, -1, 0) + dmin1(q1, 0.0) * A3(0, 0, 0) -
                                         dmax1(q2, 0.0) * A3(0, 0, 0) - dmin1(q2, 0.0) * A3(0, 1, 0));
            u[o] = nu;
            v[o] = nv;
            w[o] = nw;
        });
    } else if (c.s.convectype == 2) {
        return launch_convective_col(g, c.stream, w1, w2, w3, [=] __device__(const CvWin &W, int o) {
            const double nu = A1(0, 0, 0) +
                   tauh * (0.25 * (sq(A1(0, 0, 0) + A1(-1, 0, 0)) - sq(A1(1, 0, 0) + A1(0, 0, 0)) +
                                   (A1(0, -1, 0) + A1(0, 0, 0)) * (A2(0, 0, 0) + A2(-1, 0, 0)) -
                                   (A1(0, 1, 0) + A1(0, 0, 0)) * (A2(0, 1, 0) + A2(-1, 1, 0)) +
                                   (A1(0, 0, -1) + A1(0, 0, 0)) * (A3(0, 0, 0) + A3(-1, 0, 0)) -
                                   (A1(0, 0, 1) + A1(0, 0, 0)) * (A3(0, 0, 1) + A3(-1, 0, 1))));
            const double nv = A2(0, 0, 0) +
                   tauh * (0.25 * (sq(A2(0, 0, 0) + A2(0, -1, 0)) - sq(A2(0, 1, 0) + A2(0, 0, 0)) +
                                   (A2(-1, 0, 0) + A2(0, 0, 0)) * (A1(0, 0, 0) + A1(0, -1, 0)) -
                                   (A2(1, 0, 0) + A2(0, 0, 0)) * (A1(1, 0, 0) + A1(1, -1, 0)) +
                                   (A2(0, 0, -1) + A2(0, 0, 0)) * (A3(0, 0, 0) + A3(0, -1, 0)) -
                                   (A2(0, 0, 1) + A2(0, 0, 0)) * (A3(0, 0, 1) + A3(0, -1, 1))));
            const double nw = A3(0, 0, 0) +
                   tauh * (0.25 * (sq(A3(0, 0, 0) + A3(0, 0, -1)) - sq(A3(0, 0, 1) + A3(0, 0, 0)) +
                                   (A3(-1, 0, 0) + A3(0, 0, 0)) * (A1(0, 0, 0) + A1(0, 0, -1)) -
                                   (A3(1, 0, 0) + A3(0, 0, 0)) * (A1(1, 0, 0) + A1(1, 0, -1)) +
                                   (A3(0, -1, 0) + A3(0, 0, 0)) * (A2(0, 0, 0) + A2(0, 0, -1)) -
                                   (A3(0, 1, 0) + A3(0, 0, 0)) * (A2(0, 1, 0) + A2(0, 1, -1))));
            u[o] = nu;
            v[o] = nv;
            w[o] = nw;
#undef A1
#undef A2
#undef A3
        });
    }
    return launch_interior(g, c.stream, [=] __device__(int i, int j, int k) {
        const int o = g.at(i, j, k);
        const double nu = w1[o], nv = w2[o], nw = w3[o];
        u[o] = nu;
        v[o] = nv;
        w[o] = nw;
    });
}

int op_convective(Ctx &c, const double *w1, const double *w2, const double *w3) { return convective(c, w1, w2, w3); }

static int divergence(Ctx &c, double *w3)
{
    const Grid g = c.g;
    const double *u = c.u, *v = c.v, *w = c.w;
    const double hi2 = 1.0 / (c.s.h * c.s.h);
    const int sy = g.sy, sz = g.sz;
    return launch_interior(g, c.stream, [=] __device__(int i, int j, int k) {
        const int o = g.at(i, j, k);
        w3[o] = (u[o + 1] - u[o] + v[o + sy] - v[o] + w[o + sz] - w[o]) * hi2;
    });
}

int op_solvelpres(Ctx &c, int t)
{
    Level &L = c.lev[c.ng - 1];
    double *w1 = L.u, *w2 = L.res, *w3 = L.rhs;
    if (stress(c)) return -1;
    if (op_bc_tensor(c)) return -1;
    const double r2 = c.s.rho2;
    const double r1 = .5 * (c.s.rho1 - c.s.rho2);
    // with the split FP32 V-cycle already set up at the coefficient type it
    // will use, abcei also writes the finest level's split copy, which the
    // solve's set-up then skips
    void *split = nullptr;
    int split_ct = RB_CT_F32;
    if (rb_enabled(c)) {
        const int want = rb_coef_type(c);
        if (L.rb_coef && L.rb_ct == want) {
            split = L.rb_coef;
            split_ct = want;
        }
    }
    if (mg_abcei(c.g, c.stream, L.a, L.b, L.c, L.ei, L.cc, r1, r2, split, split_ct)) return -1;
    c.top_coef_ready = split != nullptr;
    if (predictor(c, w1, w2, w3)) return -1;
    if (op_bc_w(c, w1, w2, w3)) return -1;
    if (convective(c, w1, w2, w3)) return -1;
    if (op_bc_v(c)) return -1;
    if (divergence(c, w3)) return -1;
    if (c.mg_mode == 0) {
        if (op_mglin(c)) return -1;
    } else {
        if (op_mg_solve_tol(c)) return -1;
    }
    // projection (main.f:3375-3383)
    {
        const Grid g = c.g;
        double *u = c.u, *v = c.v, *w = c.w;
        const double *a = L.a, *b = L.b, *cf = L.c;
        const int sy = g.sy, sz = g.sz;
        if (launch_interior(g, c.stream, [=] __device__(int i, int j, int k) {
                const int o = g.at(i, j, k);
                const double p0 = w1[o];
                const double nu = u[o] - (p0 - w1[o - 1]) * a[o - 1];
                const double nv = v[o] - (p0 - w1[o - sy]) * b[o - sy];
                const double nw = w[o] - (p0 - w1[o - sz]) * cf[o - sz];
                u[o] = nu;
                v[o] = nv;
                w[o] = nw;
            }))
            return -1;
    }
    if (op_bc_v(c)) return -1;
    // adaptive smoothing counts (main.f:3389-3416)
    if (t <= 1 || t % 10 == 0) {
        if (divergence(c, w3)) return -1;
        const Grid g = c.g;
        if (reduce_interior<true>(g, c.stream, c.red_partial, c.dscal + 9,
                                  [=] __device__(int i, int j, int k) { return fabs(w3[g.at(i, j, k)]); }))
            return -1;
        DG_HIP(hipMemcpyAsync(c.hscal + 9, c.dscal + 9, sizeof(double), hipMemcpyDeviceToHost, c.stream));
        DG_HIP(hipStreamSynchronize(c.stream));
        const double epsdiv = c.s.h * c.hscal[9];
        c.epsdiv_last = epsdiv;
        if (!std::isfinite(epsdiv)) {
            set_error("solvelpres: non-finite divergence after the projection (step %d)", t);
            return -1;
        }
        if (epsdiv > 25.0 && c.s.ntop < 30) {
            c.s.ntop += 2; c.s.npre += 2; c.s.npost += 2;
        } else if (epsdiv < 2.5 && c.s.ntop > 4) {
            c.s.npre -= 1; c.s.ntop -= 1; c.s.npost -= 1;
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// initial condition smooth2 (main.f:5678-5752) for the droplet case.
// xx(1), yy(1), zz(1) are never assigned by the reference (main.f:5683); the
// natural extension (1-1.5)h is used.  The 30^3 sub-cell count is turned
// into the reference's sequential sum of 1/27000 increments.

int op_smooth2(Ctx &c)
{
    const Grid g = c.g;
    Level &L = c.lev[c.ng - 1];
    double *cc = L.cc;
    Scalars &s = c.s;
    const double h = s.h;
    const double x0 = (double)(g.nx - 2) * h / 4.0;
    const double y0 = s.y0, z0 = s.z0, r2 = s.radius * s.radius;
    if (launch_interior_rr(g, c.stream, [=] __device__(int i, int j, int k) {
            const double xa = ((double)(i - 1) - 1.5) * h, xb = ((double)i - 1.5) * h;
            const double ya = ((double)(j - 1) - 1.5) * h, yb = ((double)j - 1.5) * h;
            const double za = ((double)(k - 1) - 1.5) * h, zb = ((double)k - 1.5) * h;
            const int o = g.at(i, j, k);
            double dmin = dmin1(sq(xa - x0), sq(xb - x0)) + dmin1(sq(ya - y0), sq(yb - y0)) + dmin1(sq(za - z0), sq(zb - z0));
            if (!(dmin < r2)) { cc[o] = 0.0; return; }
            double dmax = dmax1(sq(xa - x0), sq(xb - x0)) + dmax1(sq(ya - y0), sq(yb - y0)) + dmax1(sq(za - z0), sq(zb - z0));
            if (!(dmax > r2)) { cc[o] = 1.0; return; }
            const int index = 30;
            int cnt = 0;
            for (int l = 1; l <= index; l++) {
                double px = xa + ((double)l - .5) / (double)index * h - x0;
                double px2 = sq(px);
                for (int m = 1; m <= index; m++) {
                    double py = ya + ((double)m - .5) / (double)index * h - y0;
                    double pxy = px2 + sq(py);
                    for (int n = 1; n <= index; n++) {
                        double pz = za + ((double)n - .5) / (double)index * h - z0;
                        if (pxy + sq(pz) < r2) cnt++;
                    }
                }
            }
            const double inc = 1.0 / (double)(index * index * index);
            double acc = 0.0;
            for (int q = 0; q < cnt; q++) acc = acc + inc;
            cc[o] = acc;
        }))
        return -1;
    if (op_bc_c(c, cc)) return -1;
    double *u = c.u, *v = c.v, *w = c.w;
    const double yv1 = s.yvel1, yv2 = s.yvel2, zv1 = s.zvel1, zv2 = s.zvel2;
    return launch_linear(g.n(), c.stream, [=] __device__(size_t p) {
        u[p] = 0.0;
        v[p] = yv1 * cc[p] + yv2 * (1.0 - cc[p]);
        w[p] = zv1 * cc[p] + zv2 * (1.0 - cc[p]);
    });
}

// smooth2 (main.f:5678-5752) for several droplets.  params holds 7 values
// per droplet in the solver's units (lengths in units of length_SI, the
// grid's coordinates xx(i) = (i - 1.5) h of smooth2; velocities in units of
// vscale): centre x0, y0, z0, radius r, liquid velocity u, v, w.  A cell is
// full when it lies inside one droplet, empty when it meets none, and
// otherwise counts the 30^3 sub-points of smooth2 that lie in any droplet
// (a point in two droplets counts once, for the first).  Velocities follow
// smooth2's mixture rule with the velocity of the cell's droplet (the full
// cell's, or the one holding most of its points; ghost cells take that of
// the nearest interior cell): v = v_d cc + yvel2 (1 - cc), w likewise, and
// u = u_d cc (smooth2 leaves u = 0, main.f:5741).  With one droplet at
// smooth2's centre and (0, yvel1, zvel1) this is smooth2, bit for bit.
constexpr int DROP_MAX_CANDIDATES = 8;

int op_droplets(Ctx &c, int n, const double *params)
{
    const Grid g = c.g;
    Level &L = c.lev[c.ng - 1];
    double *cc = L.cc;
    const double h = c.s.h;
    double *dp = nullptr;
    int *owner = nullptr, *flag = nullptr;
    DG_HIP(hipMalloc(&dp, sizeof(double) * 7 * (size_t)n));
    if (hipMalloc(&owner, sizeof(int) * g.n()) != hipSuccess || hipMalloc(&flag, sizeof(int)) != hipSuccess) {
        (void)hipFree(dp);
        if (owner) (void)hipFree(owner);
        set_error("droplets: out of device memory for the owner map");
        return -1;
    }
    int rc = 0;
    auto fail = [&](const char *what) {
        if (!rc) set_error("droplets: %s", what);
        rc = -1;
    };
    if (hipMemcpy(dp, params, sizeof(double) * 7 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemsetAsync(flag, 0, sizeof(int), c.stream) != hipSuccess)
        fail("parameter upload");
    if (!rc && launch_interior_rr(g, c.stream, [=] __device__(int i, int j, int k) {
            const double xa = ((double)(i - 1) - 1.5) * h, xb = ((double)i - 1.5) * h;
            const double ya = ((double)(j - 1) - 1.5) * h, yb = ((double)j - 1.5) * h;
            const double za = ((double)(k - 1) - 1.5) * h, zb = ((double)k - 1.5) * h;
            const int o = g.at(i, j, k);
            int cand[DROP_MAX_CANDIDATES];
            int nc = 0;
            for (int d = 0; d < n; d++) {
                const double x0 = dp[7 * d], y0 = dp[7 * d + 1], z0 = dp[7 * d + 2], r2 = dp[7 * d + 3] * dp[7 * d + 3];
                const double dmin =
                    dmin1(sq(xa - x0), sq(xb - x0)) + dmin1(sq(ya - y0), sq(yb - y0)) + dmin1(sq(za - z0), sq(zb - z0));
                if (!(dmin < r2)) continue;
                const double dmax =
                    dmax1(sq(xa - x0), sq(xb - x0)) + dmax1(sq(ya - y0), sq(yb - y0)) + dmax1(sq(za - z0), sq(zb - z0));
                if (!(dmax > r2)) {
                    cc[o] = 1.0;
                    owner[o] = d;
                    return;
                }
                if (nc == DROP_MAX_CANDIDATES) {
                    atomicOr(flag, 1);
                    continue;
                }
                cand[nc++] = d;
            }
            if (nc == 0) {
                cc[o] = 0.0;
                owner[o] = -1;
                return;
            }
            const int index = 30;
            int cnt[DROP_MAX_CANDIDATES];
            for (int q = 0; q < nc; q++) cnt[q] = 0;
            for (int l = 1; l <= index; l++) {
                const double tx = xa + ((double)l - .5) / (double)index * h;
                for (int m = 1; m <= index; m++) {
                    const double ty = ya + ((double)m - .5) / (double)index * h;
                    for (int q3 = 1; q3 <= index; q3++) {
                        const double tz = za + ((double)q3 - .5) / (double)index * h;
                        for (int q = 0; q < nc; q++) {
                            const int d = cand[q];
                            const double px = tx - dp[7 * d], py = ty - dp[7 * d + 1], pz = tz - dp[7 * d + 2];
                            if (sq(px) + sq(py) + sq(pz) < dp[7 * d + 3] * dp[7 * d + 3]) {
                                cnt[q]++;
                                break;
                            }
                        }
                    }
                }
            }
            int total = 0, best = 0;
            for (int q = 0; q < nc; q++) {
                total += cnt[q];
                if (cnt[q] > cnt[best]) best = q;
            }
            const double inc = 1.0 / (double)(index * index * index);
            double acc = 0.0;
            for (int q = 0; q < total; q++) acc = acc + inc;
            cc[o] = acc;
            owner[o] = cand[best];
        }))
        fail("launch");
    int hflag = 0;
    if (!rc && hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, c.stream) != hipSuccess) fail("flag");
    if (!rc && hipStreamSynchronize(c.stream) != hipSuccess) fail("synchronisation");
    if (!rc && hflag)
        fail("more than 8 droplets reach into one cell");
    if (!rc && op_bc_c(c, cc)) rc = -1;
    double *u = c.u, *v = c.v, *w = c.w;
    const double yv2 = c.s.yvel2, zv2 = c.s.zvel2;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (!rc && launch_linear(g.n(), c.stream, [=] __device__(size_t p) {
            const int i = (int)(p % (size_t)nx) + 1, j = (int)((p / (size_t)nx) % (size_t)ny) + 1,
                      k = (int)(p / ((size_t)nx * ny)) + 1;
            const int ii = i < 2 ? 2 : (i > nx - 1 ? nx - 1 : i), jj = j < 2 ? 2 : (j > ny - 1 ? ny - 1 : j),
                      kk = k < 2 ? 2 : (k > nz - 1 ? nz - 1 : k);
            const int d = owner[g.at(ii, jj, kk)];
            const double ud = d >= 0 ? dp[7 * d + 4] : 0.0, vd = d >= 0 ? dp[7 * d + 5] : 0.0,
                         wd = d >= 0 ? dp[7 * d + 6] : 0.0;
            const double cv = cc[p];
            u[p] = ud == 0.0 ? 0.0 : ud * cv;
            v[p] = vd * cv + yv2 * (1.0 - cv);
            w[p] = wd * cv + zv2 * (1.0 - cv);
        }))
        rc = -1;
    if (hipStreamSynchronize(c.stream) != hipSuccess) fail("synchronisation");
    (void)hipFree(dp);
    (void)hipFree(owner);
    (void)hipFree(flag);
    return rc;
}

}  // namespace dg
