/*
 * droplet_oracle.c -- CPU restatement of the per-time-step path of
 * farzan-t/3D_Droplet_simulation_2007 (main.f), written in plain C.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker for the GPU product and the
 * CPU baseline of bench.py.  The product library never links or calls it.
 *
 * Conventions (restating the reference, not improving it):
 *  - FP64 everywhere, evaluation order of every sum follows the Fortran
 *    source left to right; build with -ffp-contract=off so no FMA is formed.
 *  - REAL*4 literals inside FP64 expressions are kept as (double)(float)
 *    values: 1.e-30 (main.f:927), 0.8 (main.f:385,2162), the filterc27
 *    weights (main.f:5479-5489).
 *  - Scratch aliasing of the multigrid heap is reproduced: vof1/2/3 live in
 *    the finest ei/res/rhs slots (main.f:806-816), csf's w11/w22/w33 and
 *    solvelpres's w1/w2/w3 in the finest u/res/rhs slots (main.f:52,68-69).
 *  - Ghost fills run in the reference's x -> y -> z phase order.
 *
 * Parity status of each function is recorded in the test files; the whole
 * path is compared against the flang-built reference (oracle/_ref) and the
 * committed golden fixtures under tests/golden/.
 */
#include "droplet_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* 1-based Fortran indexing into an (nx,ny,nz) column-major array */
#define AT(arr, i, j, k) (arr)[((size_t)(i) - 1) + (size_t)nx * (((size_t)(j) - 1) + (size_t)ny * ((size_t)(k) - 1))]

static double dmax1(double a, double b) { return a > b ? a : b; }
static double dmin1(double a, double b) { return a < b ? a : b; }
static double sq(double a) { return a * a; }
static double cube(double a) { return a * a * a; }

/* REAL*4 literals promoted to REAL*8 */
#define F4(x) ((double)(float)(x))

/* ------------------------------------------------------------------ */
/* utilities: fillzero main.f:1286, copy main.f:4032                    */
static void fillzero(double *a, int nx, int ny, int nz) { memset(a, 0, sizeof(double) * (size_t)nx * ny * nz); }
static void copyarr(double *dst, const double *src, int nx, int ny, int nz)
{
    memcpy(dst, src, sizeof(double) * (size_t)nx * ny * nz);
}

/* ------------------------------------------------------------------ */
/* boundary conditions, main.f:4053-4539                               */

/* bc_scalar, main.f:4298-4329 */
void or_bc_scalar(double *s, int nx, int ny, int nz)
{
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++) {
            AT(s, 1, j, k) = AT(s, 2, j, k);
            AT(s, nx, j, k) = 0.0;
        }
    for (int k = 2; k <= nz - 1; k++)
        for (int i = 1; i <= nx; i++) {
            AT(s, i, 1, k) = AT(s, i, 2, k);
            AT(s, i, ny, k) = AT(s, i, ny - 1, k);
        }
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
            AT(s, i, j, 1) = AT(s, i, j, 2);
            AT(s, i, j, nz) = AT(s, i, j, nz - 1);
        }
}

/* bc_out, main.f:4399-4430 */
static void bc_out(double *s, int nx, int ny, int nz)
{
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++) {
            AT(s, 1, j, k) = AT(s, 2, j, k);
            AT(s, nx, j, k) = AT(s, nx - 1, j, k);
        }
    for (int k = 2; k <= nz - 1; k++)
        for (int i = 1; i <= nx; i++) {
            AT(s, i, 1, k) = AT(s, i, 2, k);
            AT(s, i, ny, k) = AT(s, i, ny - 1, k);
        }
    for (int j = 1; j <= ny; j++)
        for (int i = 2; i <= nx - 1; i++) {
            AT(s, i, j, 1) = AT(s, i, j, 2);
            AT(s, i, j, nz) = AT(s, i, j, nz - 1);
        }
}

/* bc_resid, main.f:4432-4463 */
static void bc_resid(double *s, int nx, int ny, int nz)
{
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++) {
            AT(s, 1, j, k) = AT(s, 2, j, k);
            AT(s, nx, j, k) = AT(s, nx - 1, j, k);
        }
    for (int k = 2; k <= nz - 1; k++)
        for (int i = 2; i <= nx - 1; i++) {
            AT(s, i, 1, k) = 0.0;
            AT(s, i, ny, k) = 0.0;
        }
    for (int j = 1; j <= ny; j++)
        for (int i = 2; i <= nx - 1; i++) {
            AT(s, i, j, 1) = 0.0;
            AT(s, i, j, nz) = 0.0;
        }
}

/* bc_coef, main.f:4465-4494 */
static void bc_coef(double *a, double *b, double *c, const double *cc, double r1, double r2, int nx, int ny, int nz)
{
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++) {
            AT(a, 1, j, k) = AT(a, 2, j, k);
            AT(a, nx, j, k) = AT(a, nx - 1, j, k);
        }
    for (int k = 2; k <= nz - 1; k++)
        for (int i = 2; i <= nx - 1; i++)
            AT(b, i, 1, k) = 1.0 / (r1 * (AT(cc, i, 1, k) + AT(cc, i, 2, k)) + r2);
    for (int j = 1; j <= ny; j++)
        for (int i = 2; i <= nx - 1; i++)
            AT(c, i, j, 1) = 1.0 / (r1 * (AT(cc, i, j, 1) + AT(cc, i, j, 2)) + r2);
}

/* bc_flux, main.f:4496-4539 */
static void bc_flux(double *vof1, double *vof3, int indx, int nx, int ny, int nz)
{
    if (indx == 1) {
        for (int k = 2; k <= nz - 1; k++)
            for (int j = 2; j <= ny - 1; j++) {
                AT(vof3, 1, j, k) = 0.0;
                AT(vof1, nx, j, k) = 0.0;
            }
    } else if (indx == 2) {
        for (int k = 2; k <= nz - 1; k++)
            for (int i = 2; i <= nx - 1; i++) {
                AT(vof3, i, 1, k) = 0.0;
                AT(vof1, i, ny, k) = 0.0;
            }
    } else {
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++) {
                AT(vof3, i, j, 1) = 0.0;
                AT(vof1, i, j, nz) = 0.0;
            }
    }
}

/* filterc27, main.f:5473-5511 (REAL*4 weights) */
void or_filterc27(const double *cc, double *cf, int nx, int ny, int nz)
{
    const double w0 = F4(0.2729), w1 = F4(0.08185), w2 = F4(0.0178), w3 = F4(.0028);
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++) {
                double t6 = AT(cc, i - 1, j, k) + AT(cc, i + 1, j, k) + AT(cc, i, j - 1, k) + AT(cc, i, j + 1, k) +
                            AT(cc, i, j, k + 1) + AT(cc, i, j, k - 1);
                double t12 = AT(cc, i - 1, j - 1, k) + AT(cc, i - 1, j + 1, k) + AT(cc, i + 1, j - 1, k) +
                             AT(cc, i + 1, j + 1, k) + AT(cc, i, j + 1, k + 1) + AT(cc, i, j - 1, k + 1) +
                             AT(cc, i + 1, j, k + 1) + AT(cc, i - 1, j, k + 1) + AT(cc, i, j + 1, k - 1) +
                             AT(cc, i, j - 1, k - 1) + AT(cc, i + 1, j, k - 1) + AT(cc, i - 1, j, k - 1);
                double t8 = AT(cc, i + 1, j + 1, k + 1) + AT(cc, i + 1, j + 1, k - 1) + AT(cc, i + 1, j - 1, k + 1) +
                            AT(cc, i + 1, j - 1, k - 1) + AT(cc, i - 1, j + 1, k + 1) + AT(cc, i - 1, j + 1, k - 1) +
                            AT(cc, i - 1, j - 1, k + 1) + AT(cc, i - 1, j - 1, k - 1);
                double v = w0 * AT(cc, i, j, k) + w1 * t6 + w2 * t12 + w3 * t8;
                if (v > 1.0) v = 1.0;
                AT(cf, i, j, k) = v;
            }
}

/* bc_c, main.f:4331-4396.  The 27-point filter and contact-angle work on
 * temporaries at main.f:4347-4375 has no effect on any output (its result
 * mmy is never used), so it is not restated. */
static void bc_c(or_state *st, double *cc)
{
    int nx = st->nx, ny = st->ny, nz = st->nz;
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++) {
            AT(cc, 2, j, k) = AT(cc, 1, j, k);
            AT(cc, nx, j, k) = AT(cc, nx - 1, j, k);
        }
    for (int k = 2; k <= nz - 1; k++)
        for (int i = 2; i <= nx - 1; i++) {
            AT(cc, i, 1, k) = AT(cc, i, 2, k);
            if (AT(cc, i, 1, k) > 1.0) AT(cc, i, 1, k) = 1.0;
            if (AT(cc, i, 1, k) < 0.0) AT(cc, i, 1, k) = 0.0;
            AT(cc, i, ny, k) = AT(cc, i, ny - 1, k);
        }
    for (int j = 1; j <= ny; j++)
        for (int i = 2; i <= nx - 1; i++) {
            AT(cc, i, j, 1) = AT(cc, i, j, 2);
            AT(cc, i, j, nz) = AT(cc, i, j, nz - 1);
        }
}

/* bc_v, main.f:4064-4141 */
static void bc_v(or_state *st)
{
    int nx = st->nx, ny = st->ny, nz = st->nz;
    double *u = st->u, *v = st->v, *w = st->w;
    for (int k = 1; k <= nz; k++)
        for (int j = ny; j >= 1; j--) {
            double q = (j * 1.0 - 1.5) / (ny * 1.0 - 1.5) - 1.0;
            AT(u, 1, j, k) = st->xvel2 * (-(q * q) + 1.0);
            if (j == 1) AT(u, 1, 1, k) = -AT(u, 1, 2, k);
            AT(v, nx, j, k) = AT(v, nx - 1, j, k);
            AT(w, nx, j, k) = AT(w, nx - 1, j, k);
        }
    for (int k = 2; k <= nz - 1; k++)
        for (int i = 2; i <= nx - 1; i++) {
            AT(u, i, 1, k) = -AT(u, i, 2, k);
            AT(u, i, ny, k) = AT(u, i, ny - 1, k);
            AT(v, i, 2, k) = 0.0;
            AT(v, i, 1, k) = 0.0;
            AT(v, i, ny, k) = 0.0;
            AT(w, i, 1, k) = -AT(w, i, 2, k);
            AT(w, i, ny, k) = AT(w, i, ny - 1, k);
        }
    for (int j = 1; j <= ny; j++)
        for (int i = 2; i <= nx - 1; i++) {
            AT(u, i, j, 1) = AT(u, i, j, 2);
            AT(u, i, j, nz) = AT(u, i, j, nz - 1);
            AT(v, i, j, 1) = AT(v, i, j, 2);
            AT(v, i, j, nz) = AT(v, i, j, nz - 1);
            AT(w, i, j, 2) = 0.0;
            AT(w, i, j, nz) = 0.0;
        }
    double outlet = 0.0, outlet2 = 0.0;
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++) {
            AT(u, 2, j, k) = AT(u, 1, j, k);
            outlet = outlet + AT(u, 1, j, k);
            outlet2 = outlet2 + AT(u, nx - 1, j, k);
        }
    double vout = outlet / (double)(ny - 2) / (double)(nz - 2);
    double vout2 = outlet2 / (double)(ny - 2) / (double)(nz - 2);
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++) AT(u, nx, j, k) = AT(u, nx - 1, j, k) + vout - vout2;
}

/* bc_w, main.f:4143-4220 */
static void bc_w(or_state *st, double *w1, double *w2, double *w3)
{
    int nx = st->nx, ny = st->ny, nz = st->nz;
    double *u = st->u, *v = st->v, *w = st->w;
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++) {
            AT(w2, 1, j, k) = AT(v, 1, j, k);
            AT(w2, 2, j, k) = AT(w2, 1, j, k);
            AT(w2, nx, j, k) = AT(v, nx, j, k);
            AT(w3, 1, j, k) = AT(w, 1, j, k);
            AT(w3, 2, j, k) = AT(w3, 1, j, k);
            AT(w3, nx, j, k) = AT(w, nx, j, k);
        }
    for (int k = 2; k <= nz - 1; k++)
        for (int i = 2; i <= nx - 1; i++) {
            AT(w1, i, 1, k) = -AT(w1, i, 2, k);
            AT(w1, i, ny, k) = AT(w1, i, ny - 1, k);
            AT(w2, i, 1, k) = 0.0;
            AT(w2, i, 2, k) = 0.0;
            AT(w2, i, ny, k) = AT(v, i, ny, k);
            AT(w3, i, 1, k) = -AT(w3, i, 2, k);
            AT(w3, i, ny, k) = AT(w3, i, ny - 1, k);
        }
    for (int j = 1; j <= ny; j++)
        for (int i = 2; i <= nx - 1; i++) {
            AT(w1, i, j, 1) = AT(w1, i, j, 2);
            AT(w1, i, j, nz) = AT(w1, i, j, nz - 1);
            AT(w2, i, j, 1) = AT(w2, i, j, 2);
            AT(w2, i, j, nz) = AT(w2, i, j, nz - 1);
            AT(w3, i, j, 1) = AT(w, i, j, 1);
            AT(w3, i, j, 2) = AT(w3, i, j, 1);
            AT(w3, i, j, nz) = AT(w, i, j, nz);
        }
    double outlet = 0.0, outlet2 = 0.0;
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++) {
            AT(w1, 1, j, k) = AT(u, 1, j, k);
            AT(w1, 2, j, k) = AT(w1, 1, j, k);
            outlet = outlet + AT(u, 1, j, k);
            outlet2 = outlet2 + AT(w1, nx - 1, j, k);
        }
    double vout = outlet / (double)(ny - 2) / (double)(nz - 2);
    double vout2 = outlet2 / (double)(ny - 2) / (double)(nz - 2);
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++) AT(w1, nx, j, k) = AT(w1, nx - 1, j, k) + vout - vout2;
}

/* bc_tensor, main.f:4222-4296 */
static void bc_tensor(or_state *st)
{
    int nx = st->nx, ny = st->ny, nz = st->nz;
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++) {
            AT(st->s11, 1, j, k) = 0.0;
            AT(st->s12, nx, j, k) = 0.0;
            AT(st->s13, nx, j, k) = 0.0;
        }
    for (int k = 2; k <= nz - 1; k++)
        for (int i = 1; i <= nx; i++) {
            AT(st->s22, i, 1, k) = AT(st->s22, i, 2, k);
            AT(st->s12, i, ny, k) = 0.0;
            AT(st->s23, i, ny, k) = 0.0;
        }
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
            AT(st->s33, i, j, 1) = 0.0;
            AT(st->s13, i, j, 2) = 0.0;
            AT(st->s13, i, j, nz) = 0.0;
            AT(st->s23, i, j, 2) = 0.0;
            AT(st->s23, i, j, nz) = 0.0;
        }
}

/* ------------------------------------------------------------------ */
/* PLIC functions AL3D main.f:4562-4634, FL3D main.f:4639-4698          */

double or_al3d(double b1, double b2, double b3, double cc)
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

double or_fl3d(double m1, double m2, double m3, double alpha, double r0, double dr0)
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

/* ------------------------------------------------------------------ */
/* VOF: swp main.f:848-1000, VOFmethod main.f:795-824                   */

static void swp(or_state *st, const double *us, double *cc, double *vof1, double *vof2, double *vof3, int d)
{
    int nx = st->nx, ny = st->ny, nz = st->nz;
    const double tauh = st->tauh;
    const double tiny = F4(1.e-30);
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++) {
                double a1 = AT(us, i, j, k) * tauh, a2;
                if (d == 1) a2 = AT(us, i + 1, j, k) * tauh;
                else if (d == 2) a2 = AT(us, i, j + 1, k) * tauh;
                else a2 = AT(us, i, j, k + 1) * tauh;
                AT(vof1, i, j, k) = 0.0;
                AT(vof2, i, j, k) = 0.0;
                AT(vof3, i, j, k) = 0.0;
                double c0 = AT(cc, i, j, k);
                if (c0 == 1.0) {
                    AT(vof1, i, j, k) = dmax1(-a1, 0.0);
                    AT(vof2, i, j, k) = 1.0 - dmax1(a1, 0.0) + dmin1(a2, 0.0);
                    AT(vof3, i, j, k) = dmax1(a2, 0.0);
                } else if (c0 > 0.0) {
                    double mm1, mm2, mx, my, mz;
                    mm1 = AT(cc, i - 1, j - 1, k - 1) + AT(cc, i - 1, j - 1, k + 1) + AT(cc, i - 1, j + 1, k - 1) +
                          AT(cc, i - 1, j + 1, k + 1) +
                          2.0 * (AT(cc, i - 1, j - 1, k) + AT(cc, i - 1, j + 1, k) + AT(cc, i - 1, j, k - 1) +
                                 AT(cc, i - 1, j, k + 1)) +
                          4.0 * AT(cc, i - 1, j, k);
                    mm2 = AT(cc, i + 1, j - 1, k - 1) + AT(cc, i + 1, j - 1, k + 1) + AT(cc, i + 1, j + 1, k - 1) +
                          AT(cc, i + 1, j + 1, k + 1) +
                          2.0 * (AT(cc, i + 1, j - 1, k) + AT(cc, i + 1, j + 1, k) + AT(cc, i + 1, j, k - 1) +
                                 AT(cc, i + 1, j, k + 1)) +
                          4.0 * AT(cc, i + 1, j, k);
                    mx = mm1 - mm2;
                    mm1 = AT(cc, i - 1, j - 1, k - 1) + AT(cc, i - 1, j - 1, k + 1) + AT(cc, i + 1, j - 1, k - 1) +
                          AT(cc, i + 1, j - 1, k + 1) +
                          2.0 * (AT(cc, i - 1, j - 1, k) + AT(cc, i + 1, j - 1, k) + AT(cc, i, j - 1, k - 1) +
                                 AT(cc, i, j - 1, k + 1)) +
                          4.0 * AT(cc, i, j - 1, k);
                    mm2 = AT(cc, i - 1, j + 1, k - 1) + AT(cc, i - 1, j + 1, k + 1) + AT(cc, i + 1, j + 1, k - 1) +
                          AT(cc, i + 1, j + 1, k + 1) +
                          2.0 * (AT(cc, i - 1, j + 1, k) + AT(cc, i + 1, j + 1, k) + AT(cc, i, j + 1, k - 1) +
                                 AT(cc, i, j + 1, k + 1)) +
                          4.0 * AT(cc, i, j + 1, k);
                    my = mm1 - mm2;
                    mm1 = AT(cc, i - 1, j - 1, k - 1) + AT(cc, i - 1, j + 1, k - 1) + AT(cc, i + 1, j - 1, k - 1) +
                          AT(cc, i + 1, j + 1, k - 1) +
                          2.0 * (AT(cc, i - 1, j, k - 1) + AT(cc, i + 1, j, k - 1) + AT(cc, i, j - 1, k - 1) +
                                 AT(cc, i, j + 1, k - 1)) +
                          4.0 * AT(cc, i, j, k - 1);
                    mm2 = AT(cc, i - 1, j - 1, k + 1) + AT(cc, i - 1, j + 1, k + 1) + AT(cc, i + 1, j - 1, k + 1) +
                          AT(cc, i + 1, j + 1, k + 1) +
                          2.0 * (AT(cc, i - 1, j, k + 1) + AT(cc, i + 1, j, k + 1) + AT(cc, i, j - 1, k + 1) +
                                 AT(cc, i, j + 1, k + 1)) +
                          4.0 * AT(cc, i, j, k + 1);
                    mz = mm1 - mm2;
                    double invx = 1.0, invy = 1.0, invz = 1.0;
                    if (mx < 0.0) { mx = -mx; invx = -1.0; }
                    if (my < 0.0) { my = -my; invy = -1.0; }
                    if (mz < 0.0) { mz = -mz; invz = -1.0; }
                    if (st->vof_guard && mx + my + mz == 0.0) {
                        if (d == 1) mx = 1.0;
                        else if (d == 2) my = 1.0;
                        else mz = 1.0;
                    }
                    mm2 = mx + my + mz + tiny;
                    mx = mx / mm2;
                    my = my / mm2;
                    mz = mz / mm2;
                    double alpha = or_al3d(mx, my, mz, c0);
                    mx = invx * mx;
                    my = invy * my;
                    mz = invz * mz;
                    alpha = alpha + dmin1(0.0, mx) + dmin1(0.0, my) + dmin1(0.0, mz);
                    mm1 = dmax1(a1, 0.0);
                    mm2 = 1.0 - mm1 + dmin1(0.0, a2);
                    if (d == 1) {
                        mx = mx / (1.0 - a1 + a2);
                        alpha = alpha + mx * a1;
                        if (a1 < 0.0) AT(vof1, i, j, k) = or_fl3d(mx, my, mz, alpha, a1, -a1);
                        if (a2 > 0.0) AT(vof3, i, j, k) = or_fl3d(mx, my, mz, alpha, 1.0, a2);
                        AT(vof2, i, j, k) = or_fl3d(mx, my, mz, alpha, mm1, mm2);
                    } else if (d == 2) {
                        my = my / (1.0 - a1 + a2);
                        alpha = alpha + my * a1;
                        if (a1 < 0.0) AT(vof1, i, j, k) = or_fl3d(my, mz, mx, alpha, a1, -a1);
                        if (a2 > 0.0) AT(vof3, i, j, k) = or_fl3d(my, mz, mx, alpha, 1.0, a2);
                        AT(vof2, i, j, k) = or_fl3d(my, mz, mx, alpha, mm1, mm2);
                    } else {
                        mz = mz / (1.0 - a1 + a2);
                        alpha = alpha + mz * a1;
                        if (a1 < 0.0) AT(vof1, i, j, k) = or_fl3d(mz, mx, my, alpha, a1, -a1);
                        if (a2 > 0.0) AT(vof3, i, j, k) = or_fl3d(mz, mx, my, alpha, 1.0, a2);
                        AT(vof2, i, j, k) = or_fl3d(mz, mx, my, alpha, mm1, mm2);
                    }
                }
            }
    bc_flux(vof1, vof3, d, nx, ny, nz);
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++) {
                double v;
                if (d == 1) v = AT(vof3, i - 1, j, k) + AT(vof2, i, j, k) + AT(vof1, i + 1, j, k);
                else if (d == 2) v = AT(vof3, i, j - 1, k) + AT(vof2, i, j, k) + AT(vof1, i, j + 1, k);
                else v = AT(vof3, i, j, k - 1) + AT(vof2, i, j, k) + AT(vof1, i, j, k + 1);
                AT(cc, i, j, k) = dmax1(0.0, dmin1(1.0, v));
            }
    bc_c(st, cc);
}

/* fturbdif, main.f:1239-1267: in-place update (k innermost, j, then i) */
static void fturbdif(or_state *st, double *cc)
{
    int nx = st->nx, ny = st->ny, nz = st->nz;
    const double *mut = st->mut, *rh = st->rhon;
    double ftr = F4(1.0);
    ftr = ftr * st->tauh2 * 1.0 / 2.0;
    for (int i = 2; i <= nx - 1; i++)
        for (int j = 2; j <= ny - 1; j++)
            for (int k = 2; k <= nz - 1; k++) {
                double m0 = AT(mut, i, j, k) / AT(rh, i, j, k);
                AT(cc, i, j, k) =
                    AT(cc, i, j, k) +
                    ftr * ((m0 + AT(mut, i, j + 1, k) / AT(rh, i, j + 1, k)) * (AT(cc, i, j + 1, k) - AT(cc, i, j, k)) -
                           (m0 + AT(mut, i, j - 1, k) / AT(rh, i, j - 1, k)) * (AT(cc, i, j, k) - AT(cc, i, j - 1, k)) +
                           (m0 + AT(mut, i, j, k + 1) / AT(rh, i, j, k + 1)) * (AT(cc, i, j, k + 1) - AT(cc, i, j, k)) -
                           (m0 + AT(mut, i, j, k - 1) / AT(rh, i, j, k - 1)) * (AT(cc, i, j, k) - AT(cc, i, j, k - 1)));
            }
    bc_c(st, cc);
}

/* calcrho, main.f:2177-2190 */
void or_calcrho(or_state *st)
{
    int nx = st->nx, ny = st->ny, nz = st->nz;
    const double *cc = st->lev[st->ng - 1].cc;
    size_t n = (size_t)nx * ny * nz;
    for (size_t p = 0; p < n; p++) {
        st->rhon[p] = st->rho2 + cc[p] * (st->rho1 - st->rho2);
        st->mun[p] = st->mu2 + cc[p] * (st->mu1 - st->mu2);
    }
}

void or_vofmethod(or_state *st, int t)
{
    (void)t;
    or_level *L = &st->lev[st->ng - 1];
    st->tswap = st->tswap + 1;
    if (st->tswap > 3) st->tswap = 1;
    if (st->tswap % 3 == 0) {
        swp(st, st->w, L->cc, L->ei, L->res, L->rhs, 3);
        swp(st, st->u, L->cc, L->ei, L->res, L->rhs, 1);
        swp(st, st->v, L->cc, L->ei, L->res, L->rhs, 2);
    } else if (st->tswap % 2 == 0) {
        swp(st, st->v, L->cc, L->ei, L->res, L->rhs, 2);
        swp(st, st->w, L->cc, L->ei, L->res, L->rhs, 3);
        swp(st, st->u, L->cc, L->ei, L->res, L->rhs, 1);
    } else {
        swp(st, st->u, L->cc, L->ei, L->res, L->rhs, 1);
        swp(st, st->v, L->cc, L->ei, L->res, L->rhs, 2);
        swp(st, st->w, L->cc, L->ei, L->res, L->rhs, 3);
    }
    bc_c(st, L->cc);
    if (st->iturb != 0) fturbdif(st, L->cc);
    or_calcrho(st);
}

/* ------------------------------------------------------------------ */
/* k-epsilon, kepsmodel main.f:1888-2026 (iturb = 3)                    */
void or_kepsmodel(or_state *st, int t)
{
    int nx = st->nx, ny = st->ny, nz = st->nz;
    or_level *L = &st->lev[st->ng - 1];
    double *kn = L->u, *en = L->res;
    const double *u = st->u, *v = st->v, *w = st->w, *kt = st->ktur, *ep = st->eps, *mut = st->mut, *mun = st->mun,
                 *rh = st->rhon;
    const double sigk = 1.0, sige = 1.3, c1 = 1.44, c2 = 1.92;
    const double h = st->h, tau = st->tau;
    const double h2 = h * h, h22 = h2 * 2.0;
    for (int j = 2; j <= ny - 1; j++)
        for (int i = 2; i <= nx - 1; i++)
            for (int k = 2; k <= nz - 1; k++) {
                double conv, diff, prod, cip, cim, cjp, cjm, ckp, ckm;
                conv = (+dmax1(AT(u, i + 1, j, k), 0.0) * AT(kt, i, j, k) + dmin1(AT(u, i + 1, j, k), 0.0) * AT(kt, i + 1, j, k) -
                        dmax1(AT(u, i, j, k), 0.0) * AT(kt, i - 1, j, k) - dmin1(AT(u, i, j, k), 0.0) * AT(kt, i, j, k) +
                        dmax1(AT(v, i, j + 1, k), 0.0) * AT(kt, i, j, k) + dmin1(AT(v, i, j + 1, k), 0.0) * AT(kt, i, j + 1, k) -
                        dmax1(AT(v, i, j, k), 0.0) * AT(kt, i, j - 1, k) - dmin1(AT(v, i, j, k), 0.0) * AT(kt, i, j, k) +
                        dmax1(AT(w, i, j, k + 1), 0.0) * AT(kt, i, j, k) + dmin1(AT(w, i, j, k + 1), 0.0) * AT(kt, i, j, k + 1) -
                        dmax1(AT(w, i, j, k), 0.0) * AT(kt, i, j, k - 1) - dmin1(AT(w, i, j, k), 0.0) * AT(kt, i, j, k)) /
                       h;
                cip = AT(mun, i + 1, j, k) + AT(mun, i, j, k) + (AT(mut, i, j, k) + AT(mut, i + 1, j, k)) / sigk;
                cim = AT(mun, i - 1, j, k) + AT(mun, i, j, k) + (AT(mut, i, j, k) + AT(mut, i - 1, j, k)) / sigk;
                cjp = AT(mun, i, j + 1, k) + AT(mun, i, j, k) + (AT(mut, i, j, k) + AT(mut, i, j + 1, k)) / sigk;
                cjm = AT(mun, i, j - 1, k) + AT(mun, i, j, k) + (AT(mut, i, j, k) + AT(mut, i, j - 1, k)) / sigk;
                ckp = AT(mun, i, j, k + 1) + AT(mun, i, j, k) + (AT(mut, i, j, k) + AT(mut, i, j, k + 1)) / sigk;
                ckm = AT(mun, i, j, k - 1) + AT(mun, i, j, k) + (AT(mut, i, j, k) + AT(mut, i, j, k - 1)) / sigk;
                diff = 1.0 / (AT(rh, i, j, k) * h22) *
                       (cip * (AT(kt, i + 1, j, k) - AT(kt, i, j, k)) - cim * (AT(kt, i, j, k) - AT(kt, i - 1, j, k)) +
                        cjp * (AT(kt, i, j + 1, k) - AT(kt, i, j, k)) - cjm * (AT(kt, i, j, k) - AT(kt, i, j - 1, k)) +
                        ckp * (AT(kt, i, j, k + 1) - AT(kt, i, j, k)) - ckm * (AT(kt, i, j, k) - AT(kt, i, j, k - 1)));
                double dudx = AT(u, i + 1, j, k) - AT(u, i, j, k);
                double dudy = 0.25 * (AT(u, i, j, k) - AT(u, i, j - 1, k) + AT(u, i + 1, j, k) - AT(u, i + 1, j - 1, k) +
                                      AT(u, i, j + 1, k) - AT(u, i, j, k) + AT(u, i + 1, j + 1, k) - AT(u, i + 1, j, k));
                double dudz = 0.25 * (AT(u, i, j, k) - AT(u, i, j, k - 1) + AT(u, i + 1, j, k) - AT(u, i + 1, j, k - 1) +
                                      AT(u, i, j, k + 1) - AT(u, i, j, k) + AT(u, i + 1, j, k + 1) - AT(u, i + 1, j, k));
                double dvdx = 0.25 * (AT(v, i, j, k) - AT(v, i - 1, j, k) + AT(v, i + 1, j, k) - AT(v, i, j, k) +
                                      AT(v, i, j + 1, k) - AT(v, i - 1, j + 1, k) + AT(v, i + 1, j + 1, k) - AT(v, i, j + 1, k));
                double dvdy = AT(v, i, j + 1, k) - AT(v, i, j, k);
                double dvdz = 0.25 * (AT(v, i, j, k) - AT(v, i, j, k - 1) + AT(v, i, j, k + 1) - AT(v, i, j, k) +
                                      AT(v, i, j + 1, k) - AT(v, i, j + 1, k - 1) + AT(v, i, j + 1, k + 1) - AT(v, i, j + 1, k));
                double dwdx = 0.25 * (AT(w, i, j, k) - AT(w, i - 1, j, k) + AT(w, i + 1, j, k) - AT(w, i, j, k) +
                                      AT(w, i, j, k + 1) - AT(w, i - 1, j, k + 1) + AT(w, i + 1, j, k + 1) - AT(w, i, j, k + 1));
                double dwdy = 0.25 * (AT(w, i, j, k) - AT(w, i, j - 1, k) + AT(w, i, j, k + 1) - AT(w, i, j - 1, k + 1) +
                                      AT(w, i, j + 1, k) - AT(w, i, j, k) + AT(w, i, j + 1, k + 1) - AT(w, i, j, k + 1));
                double dwdz = AT(w, i, j, k + 1) - AT(w, i, j, k);
                if (t == 0 || st->iturb == 3)
                    prod = AT(mut, i, j, k) / AT(rh, i, j, k) / h2 *
                           (2.0 * (dudx * dudx + dvdy * dvdy + dwdz * dwdz + dvdx * dudy + dwdx * dudz + dvdz * dwdy) +
                            dudy * dudy + dudz * dudz + dvdx * dvdx + dvdz * dvdz + dwdx * dwdx + dwdy * dwdy);
                else /* iturb = 4: production from the Reynolds stresses (main.f:1961-1964) */
                    prod = -(AT(st->u2p, i, j, k) * dudx + AT(st->v2p, i, j, k) * dvdy + AT(st->w2p, i, j, k) * dwdz +
                             AT(st->uvp, i, j, k) * (dudy + dvdx) + AT(st->uwp, i, j, k) * (dudz + dwdx) +
                             AT(st->vwp, i, j, k) * (dvdz + dwdy)) /
                           h;
                AT(kn, i, j, k) = AT(kt, i, j, k) + tau * (-conv + diff + prod - AT(ep, i, j, k));
                conv = (+dmax1(AT(u, i + 1, j, k), 0.0) * AT(ep, i, j, k) + dmin1(AT(u, i + 1, j, k), 0.0) * AT(ep, i + 1, j, k) -
                        dmax1(AT(u, i, j, k), 0.0) * AT(ep, i - 1, j, k) - dmin1(AT(u, i, j, k), 0.0) * AT(ep, i, j, k) +
                        dmax1(AT(v, i, j + 1, k), 0.0) * AT(ep, i, j, k) + dmin1(AT(v, i, j + 1, k), 0.0) * AT(ep, i, j + 1, k) -
                        dmax1(AT(v, i, j, k), 0.0) * AT(ep, i, j - 1, k) - dmin1(AT(v, i, j, k), 0.0) * AT(ep, i, j, k) +
                        dmax1(AT(w, i, j, k + 1), 0.0) * AT(ep, i, j, k) + dmin1(AT(w, i, j, k + 1), 0.0) * AT(ep, i, j, k + 1) -
                        dmax1(AT(w, i, j, k), 0.0) * AT(ep, i, j, k - 1) - dmin1(AT(w, i, j, k), 0.0) * AT(ep, i, j, k)) /
                       h;
                cip = AT(mun, i + 1, j, k) + AT(mun, i, j, k) + (AT(mut, i, j, k) + AT(mut, i + 1, j, k)) / sige;
                cim = AT(mun, i - 1, j, k) + AT(mun, i, j, k) + (AT(mut, i, j, k) + AT(mut, i - 1, j, k)) / sige;
                cjp = AT(mun, i, j + 1, k) + AT(mun, i, j, k) + (AT(mut, i, j, k) + AT(mut, i, j + 1, k)) / sige;
                cjm = AT(mun, i, j - 1, k) + AT(mun, i, j, k) + (AT(mut, i, j, k) + AT(mut, i, j - 1, k)) / sige;
                ckp = AT(mun, i, j, k + 1) + AT(mun, i, j, k) + (AT(mut, i, j, k) + AT(mut, i, j, k + 1)) / sige;
                ckm = AT(mun, i, j, k - 1) + AT(mun, i, j, k) + (AT(mut, i, j, k) + AT(mut, i, j, k - 1)) / sige;
                diff = 1.0 / (AT(rh, i, j, k) * h22) *
                       (cip * (AT(ep, i + 1, j, k) - AT(ep, i, j, k)) - cim * (AT(ep, i, j, k) - AT(ep, i - 1, j, k)) +
                        cjp * (AT(ep, i, j + 1, k) - AT(ep, i, j, k)) - cjm * (AT(ep, i, j, k) - AT(ep, i, j - 1, k)) +
                        ckp * (AT(ep, i, j, k + 1) - AT(ep, i, j, k)) - ckm * (AT(ep, i, j, k) - AT(ep, i, j, k - 1)));
                prod = prod * c1 * AT(ep, i, j, k) / AT(kt, i, j, k);
                AT(en, i, j, k) = AT(ep, i, j, k) +
                                  tau * (-conv + diff + prod - c2 * AT(ep, i, j, k) * AT(ep, i, j, k) / AT(kt, i, j, k));
            }
    for (int j = 2; j <= ny - 1; j++)
        for (int i = 2; i <= nx - 1; i++)
            for (int k = 2; k <= nz - 1; k++) {
                AT(st->ktur, i, j, k) = AT(kn, i, j, k);
                AT(st->eps, i, j, k) = AT(en, i, j, k);
                AT(st->mut, i, j, k) =
                    AT(st->cmu, i, j, k) * AT(rh, i, j, k) * AT(st->ktur, i, j, k) * AT(st->ktur, i, j, k) / AT(st->eps, i, j, k);
            }
    or_bc_scalar(st->ktur, nx, ny, nz);
    or_bc_scalar(st->eps, nx, ny, nz);
    or_bc_scalar(st->mut, nx, ny, nz);
    if (st->iturb == 4) or_uv2p(st); /* main.f:2023 */
}

/* eddyvisc, main.f:1828-1885 (iturb = 1, 2): filter u into ufilt (the
 * finest pressure slot z(iu(ng)), main.f:35) on the interior, the mixing
 * length of each x station from the filtered profile through (j1, k1), then
 * mut from the mixing length and the y and z shear of u.  j2 is left
 * uninitialised by the reference before the first station's second branch;
 * it starts at 0 here.  A centre index (j1, k1) outside the grid, which the
 * reference would read out of bounds, leaves lmix at 0. */
void or_eddyvisc(or_state *st)
{
    int nx = st->nx, ny = st->ny, nz = st->nz;
    double *uf = st->lev[st->ng - 1].u;
    const double *u = st->u;
    const double h = st->h;
    /* filter, main.f:1271-1284 */
    for (int j = 2; j <= ny - 1; j++)
        for (int i = 2; i <= nx - 1; i++)
            for (int k = 2; k <= nz - 1; k++)
                AT(uf, i, j, k) = 0.25 * AT(u, i, j, k) +
                                  0.125 * (AT(u, i, j, k - 1) + AT(u, i, j, k + 1) + AT(u, i, j - 1, k) + AT(u, i, j + 1, k)) +
                                  0.0625 * (AT(u, i, j - 1, k - 1) + AT(u, i, j - 1, k + 1) + AT(u, i, j + 1, k - 1) +
                                            AT(u, i, j + 1, k + 1));
    const double alfa = F4(0.08), cmu1 = F4(1.0);
    const int j1 = (int)(st->y0 / h), k1 = (int)(st->z0 / h);
    const int inside = j1 >= 1 && j1 <= ny && k1 >= 1 && k1 <= nz;
    int j2 = 0;
    for (int i = 1; i <= nx; i++) {
        if (!inside) {
            st->lmix[i - 1] = 0.0;
            continue;
        }
        double umax, umin;
        if (st->xvel1 > st->xvel2) {
            umax = AT(uf, i, j1, k1);
            umin = st->xvel2;
            j2 = 0;
            for (int j = 2; j <= ny - 1; j++)
                if (AT(uf, i, j, k1) - umin > F4(0.1) * (umax - umin)) {
                    j2 = j;
                    break;
                }
        } else {
            umax = st->xvel2;
            umin = AT(uf, i, j1, k1);
            for (int j = 2; j <= ny - 1; j++)
                if (AT(uf, i, j, k1) - umin <= F4(0.1) * (umax - umin)) {
                    j2 = j;
                    break;
                }
        }
        if (j2 == 2) j2 = j1;
        st->lmix[i - 1] = alfa * h * (double)abs(j2 - j1);
    }
    const double *rh = st->rhon;
    double *mut = st->mut;
    for (int i = 2; i <= nx - 1; i++)
        for (int k = 2; k <= nz - 1; k++)
            for (int j = 2; j <= ny - 1; j++) {
                const double dudy = 0.25 *
                                    (AT(u, i, j, k) - AT(u, i, j - 1, k) + AT(u, i + 1, j, k) - AT(u, i + 1, j - 1, k) +
                                     AT(u, i, j + 1, k) - AT(u, i, j, k) + AT(u, i + 1, j + 1, k) - AT(u, i + 1, j, k)) /
                                    h;
                const double dudz = 0.25 *
                                    (AT(u, i, j, k) - AT(u, i, j, k - 1) + AT(u, i + 1, j, k) - AT(u, i + 1, j, k - 1) +
                                     AT(u, i, j, k + 1) - AT(u, i, j, k) + AT(u, i + 1, j, k + 1) - AT(u, i + 1, j, k)) /
                                    h;
                const double lm = st->lmix[i - 1] + st->lmix[i];
                AT(mut, i, j, k) = cmu1 * AT(rh, i, j, k) * 0.25 * lm * lm * sqrt(dudy * dudy + dudz * dudz);
            }
    or_bc_scalar(mut, nx, ny, nz);
}

/* uv2p, main.f:2032-2155 (iturb = 4): the anisotropic (Shih-Zhu-Lumley)
 * eddy viscosity and Reynolds stresses from the strain and rotation
 * rates.  REAL literals of the reference are rounded to single precision
 * (F4): 1./3., 2./3., 1.E-30 and sqrt(6.), which the reference evaluates
 * in REAL arithmetic. */
void or_uv2p(or_state *st)
{
    int nx = st->nx, ny = st->ny, nz = st->nz;
    const double *u = st->u, *v = st->v, *w = st->w, *kt = st->ktur, *ep = st->eps, *rh = st->rhon;
    const double h = st->h;
    const double third = (double)(1.0f / 3.0f), twothird = (double)(2.0f / 3.0f), tiny = F4(1.E-30);
    const double sixsq = (double)sqrtf(6.0f);
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++) {
                const double keps = AT(kt, i, j, k) / AT(ep, i, j, k);
                const double k2eps = keps * AT(kt, i, j, k);
                const double k2eps2 = keps * keps;
                const double k3eps2 = k2eps2 * AT(kt, i, j, k);
                const double dudx = (AT(u, i + 1, j, k) - AT(u, i, j, k)) / h;
                const double dudy = 0.25 * (AT(u, i, j, k) - AT(u, i, j - 1, k) + AT(u, i + 1, j, k) - AT(u, i + 1, j - 1, k) +
                                            AT(u, i, j + 1, k) - AT(u, i, j, k) + AT(u, i + 1, j + 1, k) - AT(u, i + 1, j, k)) /
                                    h;
                const double dudz = 0.25 * (AT(u, i, j, k) - AT(u, i, j, k - 1) + AT(u, i + 1, j, k) - AT(u, i + 1, j, k - 1) +
                                            AT(u, i, j, k + 1) - AT(u, i, j, k) + AT(u, i + 1, j, k + 1) - AT(u, i + 1, j, k)) /
                                    h;
                const double dvdx = 0.25 * (AT(v, i, j, k) - AT(v, i - 1, j, k) + AT(v, i + 1, j, k) - AT(v, i, j, k) +
                                            AT(v, i, j + 1, k) - AT(v, i - 1, j + 1, k) + AT(v, i + 1, j + 1, k) - AT(v, i, j + 1, k)) /
                                    h;
                const double dvdy = (AT(v, i, j + 1, k) - AT(v, i, j, k)) / h;
                const double dvdz = 0.25 * (AT(v, i, j, k) - AT(v, i, j, k - 1) + AT(v, i, j, k + 1) - AT(v, i, j, k) +
                                            AT(v, i, j + 1, k) - AT(v, i, j + 1, k - 1) + AT(v, i, j + 1, k + 1) - AT(v, i, j + 1, k)) /
                                    h;
                const double dwdx = 0.25 * (AT(w, i, j, k) - AT(w, i - 1, j, k) + AT(w, i + 1, j, k) - AT(w, i, j, k) +
                                            AT(w, i, j, k + 1) - AT(w, i - 1, j, k + 1) + AT(w, i + 1, j, k + 1) - AT(w, i, j, k + 1)) /
                                    h;
                const double dwdy = 0.25 * (AT(w, i, j, k) - AT(w, i, j - 1, k) + AT(w, i, j, k + 1) - AT(w, i, j - 1, k + 1) +
                                            AT(w, i, j + 1, k) - AT(w, i, j, k) + AT(w, i, j + 1, k + 1) - AT(w, i, j, k + 1)) /
                                    h;
                const double dwdz = (AT(w, i, j, k + 1) - AT(w, i, j, k)) / h;
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
                AT(st->cmu, i, j, k) = cm;
                AT(st->mut, i, j, k) = cm * AT(rh, i, j, k) * AT(kt, i, j, k) * AT(kt, i, j, k) / AT(ep, i, j, k);
                double c2 = 1.0 - 9.0 * cm * cm * sstar * sstar * k2eps2;
                if (c2 < 0.0) c2 = 0.0;
                c2 = sqrt(c2) / (1.0 + 6.0 * sstar * k2eps2 * omegastar);
                AT(st->u2p, i, j, k) = twothird * AT(kt, i, j, k) - 2.0 * cm * k2eps * sstar11 +
                                       4.0 * c2 * k3eps2 * (ss12 * omeg12 + ss13 * omeg13);
                AT(st->v2p, i, j, k) = twothird * AT(kt, i, j, k) - 2.0 * cm * k2eps * sstar22 -
                                       4.0 * c2 * k3eps2 * (ss12 * omeg12 - ss23 * omeg23);
                AT(st->w2p, i, j, k) = twothird * AT(kt, i, j, k) - 2.0 * cm * k2eps * sstar33 -
                                       4.0 * c2 * k3eps2 * (ss13 * omeg13 + ss23 * omeg23);
                AT(st->uvp, i, j, k) = -(2.0 * cm * k2eps * ss12) +
                                       2.0 * c2 * k3eps2 * (-(sstar11 * omeg12) + sstar22 * omeg12 + ss13 * omeg23 + ss23 * omeg13);
                AT(st->uwp, i, j, k) = -(2.0 * cm * k2eps * ss13) +
                                       2.0 * c2 * k3eps2 * (-(sstar11 * omeg13) + sstar33 * omeg13 - ss12 * omeg23 + ss23 * omeg12);
                AT(st->vwp, i, j, k) = -(2.0 * cm * k2eps * ss23) +
                                       2.0 * c2 * k3eps2 * (-(sstar22 * omeg23) + sstar33 * omeg23 - ss12 * omeg13 - ss13 * omeg12);
            }
    or_bc_scalar(st->mut, nx, ny, nz);
    or_bc_scalar(st->u2p, nx, ny, nz);
    or_bc_scalar(st->v2p, nx, ny, nz);
    or_bc_scalar(st->w2p, nx, ny, nz);
    or_bc_scalar(st->uvp, nx, ny, nz);
    or_bc_scalar(st->uwp, nx, ny, nz);
    or_bc_scalar(st->vwp, nx, ny, nz);
}

/* ------------------------------------------------------------------ */
/* CFL main.f:4848-4881 and checktime main.f:2156-2175                  */
double or_cfl(or_state *st)
{
    int nx = st->nx, ny = st->ny, nz = st->nz;
    double c = 0.0;
    const double *arrs[3] = {st->u, st->v, st->w};
    for (int q = 0; q < 3; q++)
        for (int k = 2; k <= nz - 1; k++)
            for (int j = 2; j <= ny - 1; j++)
                for (int i = 2; i <= nx - 1; i++) c = dmax1(c, fabs(AT(arrs[q], i, j, k)));
    return c * st->tauh;
}

void or_checktime(or_state *st, int t)
{
    (void)t;
    double cfl1 = or_cfl(st);
    st->cfl_last = cfl1;
    if (cfl1 > st->maxcfl) {
        st->tau = st->tau * (st->maxcfl * F4(0.8)) / cfl1;
        st->tauh = st->tau / st->h;
        st->tauh2 = (int)(st->tau / (st->h * st->h));
    }
    if (cfl1 <= st->maxcfl * 0.25) {
        cfl1 = st->maxcfl * 0.5;
        st->tau = st->tau * (st->maxcfl * 0.5) / cfl1;
        st->tauh = st->tau / st->h;
        st->tauh2 = (int)(st->tau / (st->h * st->h));
    }
    st->time = st->time + st->tau;
}

/* ------------------------------------------------------------------ */
/* CSF surface tension, csf main.f:2363-2481                            */
void or_csf(or_state *st)
{
    int nx = st->nx, ny = st->ny, nz = st->nz;
    or_level *L = &st->lev[st->ng - 1];
    double *ccc = L->cc, *w11 = L->u, *w22 = L->res, *w33 = L->rhs;
    size_t n = (size_t)nx * ny * nz;
    double *ccf = (double *)malloc(sizeof(double) * n);
    copyarr(ccf, ccc, nx, ny, nz);
    or_filterc27(ccf, w22, nx, ny, nz);
    copyarr(ccf, w22, nx, ny, nz);
    bc_out(ccf, nx, ny, nz);
    if (st->sigma == 0.0) {
        fillzero(st->s11, nx, ny, nz); fillzero(st->s22, nx, ny, nz); fillzero(st->s33, nx, ny, nz);
        fillzero(st->s12, nx, ny, nz); fillzero(st->s13, nx, ny, nz); fillzero(st->s23, nx, ny, nz);
        free(ccf);
        return;
    }
    const double sigh = 0.5 * (st->tau * st->tau) * st->sigma / (st->h * st->h * st->h);
    const double e50 = 1.0e-50;
    for (int j = 2; j <= ny - 1; j++)
        for (int i = 2; i <= nx - 1; i++)
            for (int k = 2; k <= nz - 1; k++) {
                double nxu = sqrt(sq(AT(ccf, i + 1, j, k) - AT(ccf, i, j, k)) +
                                  0.0625 * sq(AT(ccf, i, j + 1, k) - AT(ccf, i, j - 1, k) + AT(ccf, i + 1, j + 1, k) - AT(ccf, i + 1, j - 1, k)) +
                                  0.0625 * sq(AT(ccf, i, j, k + 1) - AT(ccf, i, j, k - 1) + AT(ccf, i + 1, j, k + 1) - AT(ccf, i + 1, j, k - 1)) +
                                  e50);
                double nxd = sqrt(sq(AT(ccf, i, j, k) - AT(ccf, i - 1, j, k)) +
                                  0.0625 * sq(AT(ccf, i, j + 1, k) - AT(ccf, i, j - 1, k) + AT(ccf, i - 1, j + 1, k) - AT(ccf, i - 1, j - 1, k)) +
                                  0.0625 * sq(AT(ccf, i, j, k + 1) - AT(ccf, i, j, k - 1) + AT(ccf, i - 1, j, k + 1) - AT(ccf, i - 1, j, k - 1)) +
                                  e50);
                double nyu = sqrt(sq(AT(ccf, i, j + 1, k) - AT(ccf, i, j, k)) +
                                  0.0625 * sq(AT(ccf, i + 1, j, k) - AT(ccf, i - 1, j, k) + AT(ccf, i + 1, j + 1, k) - AT(ccf, i - 1, j + 1, k)) +
                                  0.0625 * sq(AT(ccf, i, j, k + 1) - AT(ccf, i, j, k - 1) + AT(ccf, i, j + 1, k + 1) - AT(ccf, i, j + 1, k - 1)) +
                                  e50);
                double nyd = sqrt(sq(AT(ccf, i, j, k) - AT(ccf, i, j - 1, k)) +
                                  0.0625 * sq(AT(ccf, i + 1, j, k) - AT(ccf, i - 1, j, k) + AT(ccf, i + 1, j - 1, k) - AT(ccf, i - 1, j - 1, k)) +
                                  0.0625 * sq(AT(ccf, i, j, k + 1) - AT(ccf, i, j, k - 1) + AT(ccf, i, j - 1, k + 1) - AT(ccf, i, j - 1, k - 1)) +
                                  e50);
                double nzu = sqrt(sq(AT(ccf, i, j, k + 1) - AT(ccf, i, j, k)) +
                                  0.0625 * sq(AT(ccf, i + 1, j, k) - AT(ccf, i - 1, j, k) + AT(ccf, i + 1, j, k + 1) - AT(ccf, i - 1, j, k + 1)) +
                                  0.0625 * sq(AT(ccf, i, j + 1, k) - AT(ccf, i, j - 1, k) + AT(ccf, i, j + 1, k + 1) - AT(ccf, i, j - 1, k + 1)) +
                                  e50);
                double nzd = sqrt(sq(AT(ccf, i, j, k) - AT(ccf, i, j, k - 1)) +
                                  0.0625 * sq(AT(ccf, i + 1, j, k) - AT(ccf, i - 1, j, k) + AT(ccf, i + 1, j, k - 1) - AT(ccf, i - 1, j, k - 1)) +
                                  0.0625 * sq(AT(ccf, i, j + 1, k) - AT(ccf, i, j - 1, k) + AT(ccf, i, j + 1, k - 1) - AT(ccf, i, j - 1, k - 1)) +
                                  e50);
                double d1x = (AT(ccf, i + 1, j, k) - AT(ccf, i, j, k)) / nxu;
                double d2x = (AT(ccf, i, j, k) - AT(ccf, i - 1, j, k)) / nxd;
                double d1y = (AT(ccf, i, j + 1, k) - AT(ccf, i, j, k)) / nyu;
                double d2y = (AT(ccf, i, j, k) - AT(ccf, i, j - 1, k)) / nyd;
                double d1z = (AT(ccf, i, j, k + 1) - AT(ccf, i, j, k)) / nzu;
                double d2z = (AT(ccf, i, j, k) - AT(ccf, i, j, k - 1)) / nzd;
                AT(w11, i, j, k) = -(d1x - d2x + d1y - d2y + d1z - d2z);
            }
    bc_out(w11, nx, ny, nz);
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++) {
                AT(st->s11, i, j, k) = sigh * (AT(ccf, i, j, k) - AT(ccf, i - 1, j, k)) * AT(w11, i, j, k);
                AT(st->s22, i, j, k) = sigh * (AT(ccf, i, j, k) - AT(ccf, i, j - 1, k)) * AT(w11, i, j, k);
                AT(st->s33, i, j, k) = sigh * (AT(ccf, i, j, k) - AT(ccf, i, j, k - 1)) * AT(w11, i, j, k);
            }
    copyarr(w11, st->s11, nx, ny, nz);
    copyarr(w22, st->s22, nx, ny, nz);
    copyarr(w33, st->s33, nx, ny, nz);
    bc_tensor(st);
    free(ccf);
}

/* ------------------------------------------------------------------ */
/* Interface heights, CalcH3d main.f:1312-1810.  Per interior cell: the
 * Youngs normal from the 27-point stencil (main.f:1332-1355), the cut of
 * the unit cube by the plane of that normal holding the volume
 * min(c, 1-c), classified by the sorted normal components n1 <= n2 <= n3
 * into the cases of the reference (main.f:1377-1584: two-dimensional,
 * triangle, quadrilateral A, pentagon and hexagon by Newton iteration,
 * quadrilateral B), the six face heights h1a..h3b of that cut, mirrored
 * for c > 0.5 (main.f:1590-1600) and assigned to the x/y/z faces by which
 * normal component is the smallest and the middle one (main.f:1602-1728,
 * including its use of niz's sign for the y heights in the last branch),
 * swapped for c <= 0.5 (main.f:1733-1737), clipped to [0, 1].  The
 * reference STOPs when no case matches; stop_code records that. */
static void hset(int negative, double a, double b, double *lo, double *hi)
{
    if (negative) { *lo = a; *hi = b; }
    else { *lo = b; *hi = a; }
}

void or_calch3d(or_state *st)
{
    int nx = st->nx, ny = st->ny, nz = st->nz;
    const double *cc = st->lev[st->ng - 1].cc;
    st->stop_code = 0;
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++) {
#define C(a, b, c) AT(cc, a, b, c)
                double mm1 = C(i - 1, j - 1, k - 1) + C(i - 1, j - 1, k + 1) + C(i - 1, j + 1, k - 1) + C(i - 1, j + 1, k + 1) +
                             2.0 * (C(i - 1, j - 1, k) + C(i - 1, j + 1, k) + C(i - 1, j, k - 1) + C(i - 1, j, k + 1)) +
                             4.0 * C(i - 1, j, k);
                double mm2 = C(i + 1, j - 1, k - 1) + C(i + 1, j - 1, k + 1) + C(i + 1, j + 1, k - 1) + C(i + 1, j + 1, k + 1) +
                             2.0 * (C(i + 1, j - 1, k) + C(i + 1, j + 1, k) + C(i + 1, j, k - 1) + C(i + 1, j, k + 1)) +
                             4.0 * C(i + 1, j, k);
                const double normx = mm1 - mm2;
                mm1 = C(i - 1, j - 1, k - 1) + C(i - 1, j - 1, k + 1) + C(i + 1, j - 1, k - 1) + C(i + 1, j - 1, k + 1) +
                      2.0 * (C(i - 1, j - 1, k) + C(i + 1, j - 1, k) + C(i, j - 1, k - 1) + C(i, j - 1, k + 1)) +
                      4.0 * C(i, j - 1, k);
                mm2 = C(i - 1, j + 1, k - 1) + C(i - 1, j + 1, k + 1) + C(i + 1, j + 1, k - 1) + C(i + 1, j + 1, k + 1) +
                      2.0 * (C(i - 1, j + 1, k) + C(i + 1, j + 1, k) + C(i, j + 1, k - 1) + C(i, j + 1, k + 1)) +
                      4.0 * C(i, j + 1, k);
                const double normy = mm1 - mm2;
                mm1 = C(i - 1, j - 1, k - 1) + C(i - 1, j + 1, k - 1) + C(i + 1, j - 1, k - 1) + C(i + 1, j + 1, k - 1) +
                      2.0 * (C(i - 1, j, k - 1) + C(i + 1, j, k - 1) + C(i, j - 1, k - 1) + C(i, j + 1, k - 1)) +
                      4.0 * C(i, j, k - 1);
                mm2 = C(i - 1, j - 1, k + 1) + C(i - 1, j + 1, k + 1) + C(i + 1, j - 1, k + 1) + C(i + 1, j + 1, k + 1) +
                      2.0 * (C(i - 1, j, k + 1) + C(i + 1, j, k + 1) + C(i, j - 1, k + 1) + C(i, j + 1, k + 1)) +
                      4.0 * C(i, j, k + 1);
                const double normz = mm1 - mm2;
                const double c0 = C(i, j, k);
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
                    st->stop_code = 1;  /* "NO MATCH CASES FOUND" */
                    return;
                }
                if (c0 > 0.5) {
                    h1a = 1.0 - h1a; h1b = 1.0 - h1b; h2a = 1.0 - h2a;
                    h2b = 1.0 - h2b; h3a = 1.0 - h3a; h3b = 1.0 - h3b;
                }
                double xl = 0, xh = 0, yl = 0, yh = 0, zl = 0, zh = 0;
                if (absxn == n1) {
                    hset(nix < 0, h1a, h1b, &xl, &xh);
                    if (fabs(absyn - n2) < 1.0e-6) {
                        hset(niy < 0, h2a, h2b, &yl, &yh);
                        hset(niz < 0, h3a, h3b, &zl, &zh);
                    } else if (fabs(abszn - n2) < 1.0e-6) {
                        hset(niy < 0, h3a, h3b, &yl, &yh);
                        hset(niz < 0, h2a, h2b, &zl, &zh);
                    } else {
                        st->stop_code = 2;
                        return;
                    }
                } else if (absyn == n1) {
                    hset(niy < 0, h1a, h1b, &yl, &yh);
                    if (fabs(absxn - n2) < 1.0e-6) {
                        hset(nix < 0, h2a, h2b, &xl, &xh);
                        hset(niz < 0, h3a, h3b, &zl, &zh);
                    } else if (fabs(abszn - n2) < 1.0e-6) {
                        hset(nix < 0, h3a, h3b, &xl, &xh);
                        hset(niz < 0, h2a, h2b, &zl, &zh);
                    } else {
                        st->stop_code = 2;
                        return;
                    }
                } else if (abszn == n1) {
                    hset(niz < 0, h1a, h1b, &zl, &zh);
                    if (fabs(absxn - n2) < 1.0e-6) {
                        hset(nix < 0, h2a, h2b, &xl, &xh);
                        hset(niz < 0, h3a, h3b, &yl, &yh);  /* the y heights follow niz's sign (main.f:1691-1697) */
                    } else if (fabs(absyn - n2) < 1.0e-6) {
                        hset(nix < 0, h3a, h3b, &xl, &xh);
                        hset(niz < 0, h2a, h2b, &yl, &yh);
                    } else {
                        st->stop_code = 2;
                        return;
                    }
                } else {
                    /* no component equals the minimum: a NaN normal; the
                     * reference assigns nothing and keeps the old heights */
                    xl = AT(st->hxmin, i, j, k); xh = AT(st->hxmax, i, j, k);
                    yl = AT(st->hymin, i, j, k); yh = AT(st->hymax, i, j, k);
                    zl = AT(st->hzmin, i, j, k); zh = AT(st->hzmax, i, j, k);
                }
                if (c0 <= 0.5) {
                    double t;
                    t = xl; xl = xh; xh = t;
                    t = yl; yl = yh; yh = t;
                    t = zl; zl = zh; zh = t;
                }
                double hv[6] = {xl, xh, yl, yh, zl, zh};
                for (int q = 0; q < 6; q++) if (hv[q] < 0.0) hv[q] = 0.0;
                for (int q = 0; q < 6; q++) if (hv[q] > 1.0) hv[q] = 1.0;
                AT(st->hxmin, i, j, k) = hv[0]; AT(st->hxmax, i, j, k) = hv[1];
                AT(st->hymin, i, j, k) = hv[2]; AT(st->hymax, i, j, k) = hv[3];
                AT(st->hzmin, i, j, k) = hv[4]; AT(st->hzmax, i, j, k) = hv[5];
            }
}

/* the curvature -div(grad c / |grad c|) of SGIP and SGIP2 (main.f:2878-2921,
 * 3074-3117): csf's face-normalised differences, with the tangential sums
 * of the x-down, y-up and z-down faces added in SGIP's own order */
static void sgip_curvature(const double *ccf, double *w11, int nx, int ny, int nz)
{
    const double e50 = 1.0e-50;
    for (int j = 2; j <= ny - 1; j++)
        for (int i = 2; i <= nx - 1; i++)
            for (int k = 2; k <= nz - 1; k++) {
#define F(a, b, c) AT(ccf, a, b, c)
                double nxu = sqrt(sq(F(i + 1, j, k) - F(i, j, k)) +
                                  0.0625 * sq(F(i, j + 1, k) - F(i, j - 1, k) + F(i + 1, j + 1, k) - F(i + 1, j - 1, k)) +
                                  0.0625 * sq(F(i, j, k + 1) - F(i, j, k - 1) + F(i + 1, j, k + 1) - F(i + 1, j, k - 1)) + e50);
                double nxd = sqrt(sq(F(i, j, k) - F(i - 1, j, k)) +
                                  0.0625 * sq(F(i - 1, j + 1, k) - F(i - 1, j - 1, k) + F(i, j + 1, k) - F(i, j - 1, k)) +
                                  0.0625 * sq(F(i - 1, j, k + 1) - F(i - 1, j, k - 1) + F(i, j, k + 1) - F(i, j, k - 1)) + e50);
                double nyu = sqrt(sq(F(i, j + 1, k) - F(i, j, k)) +
                                  0.0625 * sq(F(i + 1, j + 1, k) - F(i - 1, j + 1, k) + F(i + 1, j, k) - F(i - 1, j, k)) +
                                  0.0625 * sq(F(i, j + 1, k + 1) - F(i, j + 1, k - 1) + F(i, j, k + 1) - F(i, j, k - 1)) + e50);
                double nyd = sqrt(sq(F(i, j, k) - F(i, j - 1, k)) +
                                  0.0625 * sq(F(i + 1, j, k) - F(i - 1, j, k) + F(i + 1, j - 1, k) - F(i - 1, j - 1, k)) +
                                  0.0625 * sq(F(i, j, k + 1) - F(i, j, k - 1) + F(i, j - 1, k + 1) - F(i, j - 1, k - 1)) + e50);
                double nzu = sqrt(sq(F(i, j, k + 1) - F(i, j, k)) +
                                  0.0625 * sq(F(i + 1, j, k) - F(i - 1, j, k) + F(i + 1, j, k + 1) - F(i - 1, j, k + 1)) +
                                  0.0625 * sq(F(i, j + 1, k) - F(i, j - 1, k) + F(i, j + 1, k + 1) - F(i, j - 1, k + 1)) + e50);
                double nzd = sqrt(sq(F(i, j, k) - F(i, j, k - 1)) +
                                  0.0625 * sq(F(i + 1, j, k - 1) - F(i - 1, j, k - 1) + F(i + 1, j, k) - F(i - 1, j, k)) +
                                  0.0625 * sq(F(i, j + 1, k - 1) - F(i, j - 1, k - 1) + F(i, j + 1, k) - F(i, j - 1, k)) + e50);
                double d1x = (F(i + 1, j, k) - F(i, j, k)) / nxu;
                double d2x = (F(i, j, k) - F(i - 1, j, k)) / nxd;
                double d1y = (F(i, j + 1, k) - F(i, j, k)) / nyu;
                double d2y = (F(i, j, k) - F(i, j - 1, k)) / nyd;
                double d1z = (F(i, j, k + 1) - F(i, j, k)) / nzu;
                double d2z = (F(i, j, k) - F(i, j, k - 1)) / nzd;
#undef F
                AT(w11, i, j, k) = -(d1x - d2x + d1y - d2y + d1z - d2z);
            }
}

/* SIND / COSD of the reference, built with -Ddsind=sind -Ddcosd=cosd
 * (oracle/build_ref.sh): the Fortran 2023 intrinsics, the sine and cosine
 * of the argument times pi/180 */
#define OR_PI 3.14159265358979323846
static double sind_(double x) { return sin(x * (OR_PI / 180.0)); }
static double cosd_(double x) { return cos(x * (OR_PI / 180.0)); }

/* SGIP surface tension, main.f:2835-3027: one filter pass (as csf), the
 * curvature, S_dd = sigh (w11_i + w11_{i-1}) (ccf_{i+1} - ccf_{i-1})
 * |L - L_down| / 2 with L the mean interface height of CalcH3d along d;
 * on the first row above the wall (j = 2) a contact-angle form with the
 * angle theta from the contact-line velocity (main.f:2954-2999); bc_out
 * of S11, S22, S33 and the copies w = S.  No bc_tensor. */
void or_sgip(or_state *st)
{
    int nx = st->nx, ny = st->ny, nz = st->nz;
    or_level *L = &st->lev[st->ng - 1];
    double *ccc = L->cc, *w11 = L->u, *w22 = L->res, *w33 = L->rhs;
    size_t n = (size_t)nx * ny * nz;
    double *ccf = (double *)malloc(sizeof(double) * n);
    copyarr(ccf, ccc, nx, ny, nz);
    or_filterc27(ccf, w22, nx, ny, nz);
    copyarr(ccf, w22, nx, ny, nz);
    bc_out(ccf, nx, ny, nz);
    if (st->sigma == 0.0) {
        fillzero(st->s11, nx, ny, nz); fillzero(st->s22, nx, ny, nz); fillzero(st->s33, nx, ny, nz);
        fillzero(st->s12, nx, ny, nz); fillzero(st->s13, nx, ny, nz); fillzero(st->s23, nx, ny, nz);
        free(ccf);
        return;
    }
    const double sigh = 0.5 * (st->tau * st->tau) * st->sigma / (st->h * st->h * st->h);
    sgip_curvature(ccf, w11, nx, ny, nz);
    bc_out(w11, nx, ny, nz);
    const double *u = st->u, *w = st->w;
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++) {
                const double Lx = (AT(st->hxmax, i, j, k) + AT(st->hxmin, i, j, k)) / 2.0;
                const double Lxd = (AT(st->hxmax, i - 1, j, k) + AT(st->hxmin, i - 1, j, k)) / 2.0;
                const double Ly = (AT(st->hymax, i, j, k) + AT(st->hymin, i, j, k)) / 2.0;
                const double Lyd = (AT(st->hymax, i, j - 1, k) + AT(st->hymin, i, j - 1, k)) / 2.0;
                const double Lz = (AT(st->hzmax, i, j, k) + AT(st->hzmin, i, j, k)) / 2.0;
                const double Lzd = (AT(st->hzmax, i, j, k - 1) + AT(st->hzmin, i, j, k - 1)) / 2.0;
                AT(st->s11, i, j, k) = sigh * (AT(w11, i, j, k) + AT(w11, i - 1, j, k)) *
                                       (AT(ccf, i + 1, j, k) - AT(ccf, i - 1, j, k)) * fabs(Lx - Lxd) / 2.0;
                AT(st->s22, i, j, k) = sigh * (AT(w11, i, j, k) + AT(w11, i, j - 1, k)) *
                                       (AT(ccf, i, j + 1, k) - AT(ccf, i, j - 1, k)) * fabs(Ly - Lyd) / 2.0;
                AT(st->s33, i, j, k) = sigh * (AT(w11, i, j, k) + AT(w11, i, j, k - 1)) *
                                       (AT(ccf, i, j, k + 1) - AT(ccf, i, j, k - 1)) * fabs(Lz - Lzd) / 2.0;
                if (j == 2) {
                    double mx = AT(ccf, i + 1, j, k) - AT(ccf, i - 1, j, k);
                    double mz = AT(ccf, i, j, k + 1) - AT(ccf, i, j, k - 1);
                    const double mmx = mx / sqrt(mz * mz + mx * mx + 1.0e-50);
                    const double mmz = mz / sqrt(mz * mz + mx * mx + 1.0e-50);
                    double normv = mmx * AT(u, i, j, k) + mmz * AT(w, i, j, k);
                    normv = -normv;
                    double theta = 0.0;
                    if (fabs(normv) > (double)0.1f) {
                        if (normv > 0) theta = 115.0;
                        if (normv < 0) theta = 45.0;
                    } else {
                        theta = (double)350.0f * normv + 80.0;
                    }
                    const double my = AT(ccf, i, j + 1, k) - AT(ccf, i, j - 1, k);
                    const double mmy = cosd_(theta);
                    AT(st->s11, i, j, k) = sigh * (AT(w11, i, j, k) + AT(w11, i - 1, j, k)) * mx * sind_(theta) / 2.0;
                    AT(st->s22, i, j, k) = sigh * (AT(w11, i, j, k) + AT(w11, i, j - 1, k)) * mmy *
                                           sqrt(mz * mz + mx * mx + my * my) * -1.0;
                    AT(st->s33, i, j, k) = sigh * (AT(w11, i, j, k) + AT(w11, i, j, k - 1)) * mz * sind_(theta) / 2.0;
                }
            }
    bc_out(st->s11, nx, ny, nz);
    bc_out(st->s22, nx, ny, nz);
    bc_out(st->s33, nx, ny, nz);
    copyarr(w11, st->s11, nx, ny, nz);
    copyarr(w22, st->s22, nx, ny, nz);
    copyarr(w33, st->s33, nx, ny, nz);
    free(ccf);
}

/* SGIP2 (second order), main.f:3029-3166: three filter passes, sigh =
 * 0.5 tau sigma / h^2, the centred form S_dd = sigh (w11_{d+1} + w11_{d-1})
 * (ccf_{d+1} - ccf_{d-1}) (L_{d+1} - L_{d-1}) / 4; no bc_out of S, no
 * bc_tensor. */
void or_sgip2(or_state *st)
{
    int nx = st->nx, ny = st->ny, nz = st->nz;
    or_level *L = &st->lev[st->ng - 1];
    double *ccc = L->cc, *w11 = L->u, *w22 = L->res, *w33 = L->rhs;
    size_t n = (size_t)nx * ny * nz;
    double *ccf = (double *)malloc(sizeof(double) * n);
    copyarr(ccf, ccc, nx, ny, nz);
    for (int pass = 0; pass < 3; pass++) {
        or_filterc27(ccf, w22, nx, ny, nz);
        copyarr(ccf, w22, nx, ny, nz);
        bc_out(ccf, nx, ny, nz);
    }
    if (st->sigma == 0.0) {
        fillzero(st->s11, nx, ny, nz); fillzero(st->s22, nx, ny, nz); fillzero(st->s33, nx, ny, nz);
        fillzero(st->s12, nx, ny, nz); fillzero(st->s13, nx, ny, nz); fillzero(st->s23, nx, ny, nz);
        free(ccf);
        return;
    }
    const double sigh = 0.5 * st->tau * st->sigma / (st->h * st->h);
    sgip_curvature(ccf, w11, nx, ny, nz);
    bc_out(w11, nx, ny, nz);
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++) {
                const double Lxu = (AT(st->hxmax, i + 1, j, k) + AT(st->hxmin, i + 1, j, k)) / 2.0;
                const double Lxd = (AT(st->hxmax, i - 1, j, k) + AT(st->hxmin, i - 1, j, k)) / 2.0;
                const double Lyu = (AT(st->hymax, i, j + 1, k) + AT(st->hymin, i, j + 1, k)) / 2.0;
                const double Lyd = (AT(st->hymax, i, j - 1, k) + AT(st->hymin, i, j - 1, k)) / 2.0;
                const double Lzu = (AT(st->hzmax, i, j, k + 1) + AT(st->hzmin, i, j, k + 1)) / 2.0;
                const double Lzd = (AT(st->hzmax, i, j, k - 1) + AT(st->hzmin, i, j, k - 1)) / 2.0;
                AT(st->s11, i, j, k) = sigh * (AT(w11, i + 1, j, k) + AT(w11, i - 1, j, k)) *
                                       (AT(ccf, i + 1, j, k) - AT(ccf, i - 1, j, k)) * (Lxu - Lxd) * 0.25;
                AT(st->s22, i, j, k) = sigh * (AT(w11, i, j + 1, k) + AT(w11, i, j - 1, k)) *
                                       (AT(ccf, i, j + 1, k) - AT(ccf, i, j - 1, k)) * (Lyu - Lyd) * 0.25;
                AT(st->s33, i, j, k) = sigh * (AT(w11, i, j, k + 1) + AT(w11, i, j, k - 1)) *
                                       (AT(ccf, i, j, k + 1) - AT(ccf, i, j, k - 1)) * (Lzu - Lzd) * 0.25;
            }
    copyarr(w11, st->s11, nx, ny, nz);
    copyarr(w22, st->s22, nx, ny, nz);
    copyarr(w33, st->s33, nx, ny, nz);
    free(ccf);
}

/* ------------------------------------------------------------------ */
/* multigrid kernels, main.f:3765-3997                                  */

void or_rstrct(double *uc, const double *uf, int ncx, int ncy, int ncz)
{
    int nfx = 2 * ncx - 2, nfy = 2 * ncy - 2;
    for (int kc = 2; kc <= ncz - 1; kc++) {
        int kf = 2 * kc - 2;
        for (int jc = 2; jc <= ncy - 1; jc++) {
            int jf = 2 * jc - 2;
            for (int ic = 2; ic <= ncx - 1; ic++) {
                int iff = 2 * ic - 2;
#define UF(i, j, k) uf[((size_t)(i) - 1) + (size_t)nfx * (((size_t)(j) - 1) + (size_t)nfy * ((size_t)(k) - 1))]
                double s = UF(iff, jf, kf) + UF(iff + 1, jf, kf) + UF(iff, jf + 1, kf) + UF(iff + 1, jf + 1, kf) +
                           UF(iff, jf, kf + 1) + UF(iff + 1, jf, kf + 1) + UF(iff, jf + 1, kf + 1) + UF(iff + 1, jf + 1, kf + 1);
#undef UF
                int nx = ncx, ny = ncy;
                AT(uc, ic, jc, kc) = 0.125 * s;
            }
        }
    }
    or_bc_scalar(uc, ncx, ncy, ncz);
}

void or_acfromac(double *a, double *b, double *c, double *ei, const double *a2, const double *b2, const double *c2, int nx,
                 int ny, int nz)
{
    int nfx = 2 * nx - 2, nfy = 2 * ny - 2;
#define F2(arr, i, j, k) arr[((size_t)(i) - 1) + (size_t)nfx * (((size_t)(j) - 1) + (size_t)nfy * ((size_t)(k) - 1))]
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 1; i <= nx - 1; i++)
                AT(a, i, j, k) = 0.25 * (F2(a2, 2 * i - 1, 2 * j - 1, 2 * k - 1) + F2(a2, 2 * i - 1, 2 * j - 2, 2 * k - 1) +
                                         F2(a2, 2 * i - 1, 2 * j - 1, 2 * k - 2) + F2(a2, 2 * i - 1, 2 * j - 2, 2 * k - 2));
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 1; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++)
                AT(b, i, j, k) = 0.25 * (F2(b2, 2 * i - 1, 2 * j - 1, 2 * k - 1) + F2(b2, 2 * i - 2, 2 * j - 1, 2 * k - 1) +
                                         F2(b2, 2 * i - 1, 2 * j - 1, 2 * k - 2) + F2(b2, 2 * i - 2, 2 * j - 1, 2 * k - 2));
    for (int k = 1; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++)
                AT(c, i, j, k) = 0.25 * (F2(c2, 2 * i - 1, 2 * j - 1, 2 * k - 1) + F2(c2, 2 * i - 2, 2 * j - 1, 2 * k - 1) +
                                         F2(c2, 2 * i - 1, 2 * j - 2, 2 * k - 1) + F2(c2, 2 * i - 2, 2 * j - 2, 2 * k - 1));
#undef F2
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++)
                AT(ei, i, j, k) = 1.0 / (AT(a, i, j, k) + AT(a, i - 1, j, k) + AT(b, i, j, k) + AT(b, i, j - 1, k) +
                                         AT(c, i, j, k) + AT(c, i, j, k - 1));
}

void or_relax(double *u, double lengthsize, const double *rhs, const double *a, const double *b, const double *c,
              const double *ei, int nx, int ny, int nz, double omega)
{
    double h1 = lengthsize / (double)(nx - 2);
    double h2 = h1 * h1;
    int ksw = 1;
    for (int ipass = 1; ipass <= 2; ipass++) {
        or_bc_scalar(u, nx, ny, nz);
        int jsw = ksw;
        for (int k = 2; k <= nz - 1; k++) {
            int isw = jsw;
            for (int j = 2; j <= ny - 1; j++) {
                for (int i = isw + 1; i <= nx - 1; i += 2) {
                    double temp = (AT(a, i, j, k) * AT(u, i + 1, j, k) + AT(a, i - 1, j, k) * AT(u, i - 1, j, k) +
                                   AT(b, i, j, k) * AT(u, i, j + 1, k) + AT(b, i, j - 1, k) * AT(u, i, j - 1, k) +
                                   AT(c, i, j, k) * AT(u, i, j, k + 1) + AT(c, i, j, k - 1) * AT(u, i, j, k - 1) -
                                   h2 * AT(rhs, i, j, k)) *
                                  AT(ei, i, j, k);
                    AT(u, i, j, k) = omega * temp + (1.0 - omega) * AT(u, i, j, k);
                }
                isw = 3 - isw;
            }
            jsw = 3 - jsw;
        }
        ksw = 3 - ksw;
    }
    or_bc_scalar(u, nx, ny, nz);
}

void or_resid(double lengthsize, double *res, const double *u, const double *rhs, const double *a, const double *b,
              const double *c, int nx, int ny, int nz)
{
    double h1 = lengthsize / (double)(nx - 2);
    double h2i = 1.0 / (h1 * h1);
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++)
                AT(res, i, j, k) = -h2i * (AT(a, i, j, k) * (AT(u, i + 1, j, k) - AT(u, i, j, k)) +
                                           AT(a, i - 1, j, k) * (AT(u, i - 1, j, k) - AT(u, i, j, k)) +
                                           AT(b, i, j, k) * (AT(u, i, j + 1, k) - AT(u, i, j, k)) +
                                           AT(b, i, j - 1, k) * (AT(u, i, j - 1, k) - AT(u, i, j, k)) +
                                           AT(c, i, j, k) * (AT(u, i, j, k + 1) - AT(u, i, j, k)) +
                                           AT(c, i, j, k - 1) * (AT(u, i, j, k - 1) - AT(u, i, j, k))) +
                                   AT(rhs, i, j, k);
    bc_resid(res, nx, ny, nz);
}

void or_interp(double *uf, const double *uc, int nfx, int nfy, int nfz)
{
    int ncx = nfx / 2 + 1, ncy = nfy / 2 + 1, ncz = nfz / 2 + 1;
    int nx = nfx, ny = nfy;
    for (int kc = 2; kc <= ncz - 1; kc++) {
        int kf = 2 * kc - 2;
        for (int jc = 2; jc <= ncy - 1; jc++) {
            int jf = 2 * jc - 2;
            for (int ic = 2; ic <= ncx - 1; ic++) {
                int iff = 2 * ic - 2;
                double v = uc[((size_t)ic - 1) + (size_t)ncx * (((size_t)jc - 1) + (size_t)ncy * ((size_t)kc - 1))];
                AT(uf, iff, jf, kf) = v;
                AT(uf, iff, jf + 1, kf) = v;
                AT(uf, iff + 1, jf, kf) = v;
                AT(uf, iff + 1, jf + 1, kf) = v;
                AT(uf, iff, jf, kf + 1) = v;
                AT(uf, iff, jf + 1, kf + 1) = v;
                AT(uf, iff + 1, jf, kf + 1) = v;
                AT(uf, iff + 1, jf + 1, kf + 1) = v;
            }
        }
    }
    or_bc_scalar(uf, nfx, nfy, nfz);
}

void or_addint(double *uf, const double *uc, double *res, int nfx, int nfy, int nfz)
{
    or_interp(res, uc, nfx, nfy, nfz);
    size_t n = (size_t)nfx * nfy * nfz;
    for (size_t p = 0; p < n; p++) uf[p] = uf[p] + res[p];
}

void or_slvsml(double *u, const double *rhs, const double *a, const double *b, const double *c, const double *ei, int nx,
               int ny, int nz, int ntop, double omega, double lengthsize)
{
    for (int j = 1; j <= ntop; j++) or_relax(u, lengthsize, rhs, a, b, c, ei, nx, ny, nz, omega);
}

void or_abcei(double *a, double *b, double *c, double *ei, const double *cc, double r1, double r2, int nx, int ny, int nz)
{
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++) {
                AT(a, i, j, k) = 1.0 / (r1 * (AT(cc, i, j, k) + AT(cc, i + 1, j, k)) + r2);
                AT(b, i, j, k) = 1.0 / (r1 * (AT(cc, i, j, k) + AT(cc, i, j + 1, k)) + r2);
                AT(c, i, j, k) = 1.0 / (r1 * (AT(cc, i, j, k) + AT(cc, i, j, k + 1)) + r2);
            }
    bc_coef(a, b, c, cc, r1, r2, nx, ny, nz);
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++)
                AT(ei, i, j, k) = 1.0 / (AT(a, i, j, k) + AT(a, i - 1, j, k) + AT(b, i, j, k) + AT(b, i, j - 1, k) +
                                         AT(c, i, j, k) + AT(c, i, j, k - 1));
}

/* mglin, main.f:3645-3763: full-multigrid schedule, fixed cycle counts */
void or_mglin(or_state *st)
{
    if (!st->ref_grid) {  /* memoryaloc's STOP (main.f:274-322) */
        fprintf(stderr, "STOP: nn .ne. 2**ng + 2\n");
        exit(1);
    }
    or_level *lv = st->lev;
    int ng = st->ng;
    /* restrict rhs and coefficients down to the coarsest level */
    for (int g = ng - 1; g >= 1; g--) {
        or_level *C = &lv[g - 1], *F = &lv[g];
        or_rstrct(C->rhs, F->rhs, C->nx, C->ny, C->nz);
        or_acfromac(C->a, C->b, C->c, C->ei, F->a, F->b, F->c, C->nx, C->ny, C->nz);
    }
    or_level *L1 = &lv[0];
    fillzero(L1->u, L1->nx, L1->ny, L1->nz);
    /* first coarsest solve takes lengthscale, not lengthsize (main.f:3689) */
    or_slvsml(L1->u, L1->rhs, L1->a, L1->b, L1->c, L1->ei, L1->nx, L1->ny, L1->nz, st->ntop, st->omega, st->lengthscale);
    for (int j = 2; j <= ng; j++) {
        or_level *J = &lv[j - 1];
        or_interp(J->u, lv[j - 2].u, J->nx, J->ny, J->nz);
        for (int jc = 1; jc <= st->cycles; jc++) {
            for (int jj = j; jj >= 2; jj--) {
                or_level *F = &lv[jj - 1], *C = &lv[jj - 2];
                for (int q = 1; q <= st->npre; q++)
                    or_relax(F->u, st->lengthsize, F->rhs, F->a, F->b, F->c, F->ei, F->nx, F->ny, F->nz, st->omega);
                or_resid(st->lengthsize, F->res, F->u, F->rhs, F->a, F->b, F->c, F->nx, F->ny, F->nz);
                or_rstrct(C->rhs, F->res, C->nx, C->ny, C->nz);
                fillzero(C->u, C->nx, C->ny, C->nz);
            }
            or_slvsml(L1->u, L1->rhs, L1->a, L1->b, L1->c, L1->ei, L1->nx, L1->ny, L1->nz, st->ntop, st->omega,
                      st->lengthsize);
            for (int jj = 2; jj <= j; jj++) {
                or_level *F = &lv[jj - 1], *C = &lv[jj - 2];
                or_addint(F->u, C->u, F->res, F->nx, F->ny, F->nz);
                for (int q = 1; q <= st->npost; q++)
                    or_relax(F->u, st->lengthsize, F->rhs, F->a, F->b, F->c, F->ei, F->nx, F->ny, F->nz, st->omega);
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* convective, main.f:3441-3580, with the Reynolds-stress terms of
 * u2p..vwp in the reference's evaluation order (the arrays are zero unless
 * iturb = 4)                                                             */
static void convective(or_state *st, const double *w1, const double *w2, const double *w3)
{
    int nx = st->nx, ny = st->ny, nz = st->nz;
    double *u = st->u, *v = st->v, *w = st->w;
    const double tauh = st->tauh;
    if (st->convectype == 1) {
        for (int k = 2; k <= nz - 1; k++)
            for (int j = 2; j <= ny - 1; j++)
                for (int i = 2; i <= nx - 1; i++) {
                    double t1, t2, s1, s2, q1, q2;
                    t1 = .5 * (AT(w2, i, j, k) + AT(w2, i - 1, j, k));
                    t2 = .5 * (AT(w2, i, j + 1, k) + AT(w2, i - 1, j + 1, k));
                    s1 = .5 * (AT(w1, i - 1, j, k) + AT(w1, i, j, k));
                    s2 = .5 * (AT(w1, i, j, k) + AT(w1, i + 1, j, k));
                    q1 = .5 * (AT(w3, i, j, k) + AT(w3, i - 1, j, k));
                    q2 = .5 * (AT(w3, i, j, k + 1) + AT(w3, i - 1, j, k + 1));
                    AT(u, i, j, k) =
                        AT(w1, i, j, k) +
                        tauh * (+dmax1(t1, 0.0) * AT(w1, i, j - 1, k) + dmin1(t1, 0.0) * AT(w1, i, j, k) -
                                dmax1(t2, 0.0) * AT(w1, i, j, k) - dmin1(t2, 0.0) * AT(w1, i, j + 1, k) +
                                dmax1(s1, 0.0) * AT(w1, i - 1, j, k) + dmin1(s1, 0.0) * AT(w1, i, j, k) -
                                dmax1(s2, 0.0) * AT(w1, i, j, k) - dmin1(s2, 0.0) * AT(w1, i + 1, j, k) +
                                dmax1(q1, 0.0) * AT(w1, i, j, k - 1) + dmin1(q1, 0.0) * AT(w1, i, j, k) -
                                dmax1(q2, 0.0) * AT(w1, i, j, k) - dmin1(q2, 0.0) * AT(w1, i, j, k + 1) + AT(st->u2p, i - 1, j, k) - AT(st->u2p, i, j, k) +
                                0.25 * (+AT(st->uvp, i, j, k) - AT(st->uvp, i, j + 1, k) + AT(st->uvp, i, j - 1, k) - AT(st->uvp, i, j, k) + AT(st->uvp, i - 1, j, k) - AT(st->uvp, i - 1, j + 1, k) + AT(st->uvp, i - 1, j - 1, k) - AT(st->uvp, i - 1, j, k) + AT(st->uwp, i, j, k) - AT(st->uwp, i, j, k + 1) + AT(st->uwp, i, j, k - 1) - AT(st->uwp, i, j, k) + AT(st->uwp, i - 1, j, k) - AT(st->uwp, i - 1, j, k + 1) + AT(st->uwp, i - 1, j, k - 1) - AT(st->uwp, i - 1, j, k)));
                    t1 = .5 * (AT(w1, i, j, k) + AT(w1, i, j - 1, k));
                    t2 = .5 * (AT(w1, i + 1, j, k) + AT(w1, i + 1, j - 1, k));
                    s1 = .5 * (AT(w2, i, j - 1, k) + AT(w2, i, j, k));
                    s2 = .5 * (AT(w2, i, j, k) + AT(w2, i, j + 1, k));
                    q1 = .5 * (AT(w3, i, j, k) + AT(w3, i, j - 1, k));
                    q2 = .5 * (AT(w3, i, j, k + 1) + AT(w3, i, j - 1, k + 1));
                    AT(v, i, j, k) =
                        AT(w2, i, j, k) +
                        tauh * (+dmax1(t1, 0.0) * AT(w2, i - 1, j, k) + dmin1(t1, 0.0) * AT(w2, i, j, k) -
                                dmax1(t2, 0.0) * AT(w2, i, j, k) - dmin1(t2, 0.0) * AT(w2, i + 1, j, k) +
                                dmax1(s1, 0.0) * AT(w2, i, j - 1, k) + dmin1(s1, 0.0) * AT(w2, i, j, k) -
                                dmax1(s2, 0.0) * AT(w2, i, j, k) - dmin1(s2, 0.0) * AT(w2, i, j + 1, k) +
                                dmax1(q1, 0.0) * AT(w2, i, j, k - 1) + dmin1(q1, 0.0) * AT(w2, i, j, k) -
                                dmax1(q2, 0.0) * AT(w2, i, j, k) - dmin1(q2, 0.0) * AT(w2, i, j, k + 1) + AT(st->v2p, i, j - 1, k) - AT(st->v2p, i, j, k) +
                                0.25 * (+AT(st->uvp, i - 1, j, k) - AT(st->uvp, i, j, k) + AT(st->uvp, i, j, k) - AT(st->uvp, i + 1, j, k) + AT(st->uvp, i - 1, j - 1, k) - AT(st->uvp, i, j - 1, k) + AT(st->uvp, i, j - 1, k) - AT(st->uvp, i + 1, j - 1, k) + AT(st->vwp, i, j, k - 1) - AT(st->vwp, i, j, k) + AT(st->vwp, i, j, k) - AT(st->vwp, i, j, k + 1) + AT(st->vwp, i, j - 1, k - 1) - AT(st->vwp, i, j - 1, k) + AT(st->vwp, i, j - 1, k) - AT(st->vwp, i, j - 1, k + 1)));
                    t1 = .5 * (AT(w1, i, j, k) + AT(w1, i, j, k - 1));
                    t2 = .5 * (AT(w1, i + 1, j, k) + AT(w1, i + 1, j, k - 1));
                    s1 = .5 * (AT(w3, i, j, k - 1) + AT(w3, i, j, k));
                    s2 = .5 * (AT(w3, i, j, k) + AT(w3, i, j, k + 1));
                    q1 = .5 * (AT(w2, i, j, k) + AT(w2, i, j, k - 1));
                    q2 = .5 * (AT(w2, i, j + 1, k) + AT(w2, i, j + 1, k - 1));
                    AT(w, i, j, k) =
                        AT(w3, i, j, k) +
                        tauh * (+dmax1(t1, 0.0) * AT(w3, i - 1, j, k) + dmin1(t1, 0.0) * AT(w3, i, j, k) -
                                dmax1(t2, 0.0) * AT(w3, i, j, k) - dmin1(t2, 0.0) * AT(w3, i + 1, j, k) +
                                dmax1(s1, 0.0) * AT(w3, i, j, k - 1) + dmin1(s1, 0.0) * AT(w3, i, j, k) -
                                dmax1(s2, 0.0) * AT(w3, i, j, k) - dmin1(s2, 0.0) * AT(w3, i, j, k + 1) +
                                dmax1(q1, 0.0) * AT(w3, i, j - 1, k) + dmin1(q1, 0.0) * AT(w3, i, j, k) -
                                dmax1(q2, 0.0) * AT(w3, i, j, k) - dmin1(q2, 0.0) * AT(w3, i, j + 1, k) + AT(st->w2p, i, j, k - 1) - AT(st->w2p, i, j, k) +
                                0.25 * (+AT(st->uwp, i - 1, j, k) - AT(st->uwp, i, j, k) + AT(st->uwp, i, j, k) - AT(st->uwp, i + 1, j, k) + AT(st->uwp, i - 1, j, k - 1) - AT(st->uwp, i, j, k - 1) + AT(st->uwp, i, j, k - 1) - AT(st->uwp, i + 1, j, k - 1) + AT(st->vwp, i, j - 1, k) - AT(st->vwp, i, j, k) + AT(st->vwp, i, j, k) - AT(st->vwp, i, j + 1, k) + AT(st->vwp, i, j - 1, k - 1) - AT(st->vwp, i, j, k - 1) + AT(st->vwp, i, j, k - 1) - AT(st->vwp, i, j + 1, k - 1)));
                }
    } else if (st->convectype == 2) {
        for (int k = 2; k <= nz - 1; k++)
            for (int j = 2; j <= ny - 1; j++)
                for (int i = 2; i <= nx - 1; i++) {
                    AT(u, i, j, k) =
                        AT(w1, i, j, k) +
                        tauh * (0.25 * (sq(AT(w1, i, j, k) + AT(w1, i - 1, j, k)) - sq(AT(w1, i + 1, j, k) + AT(w1, i, j, k)) +
                                        (AT(w1, i, j - 1, k) + AT(w1, i, j, k)) * (AT(w2, i, j, k) + AT(w2, i - 1, j, k)) -
                                        (AT(w1, i, j + 1, k) + AT(w1, i, j, k)) * (AT(w2, i, j + 1, k) + AT(w2, i - 1, j + 1, k)) +
                                        (AT(w1, i, j, k - 1) + AT(w1, i, j, k)) * (AT(w3, i, j, k) + AT(w3, i - 1, j, k)) -
                                        (AT(w1, i, j, k + 1) + AT(w1, i, j, k)) * (AT(w3, i, j, k + 1) + AT(w3, i - 1, j, k + 1)) +
                                        AT(st->uvp, i, j, k) - AT(st->uvp, i, j + 1, k) + AT(st->uvp, i, j - 1, k) - AT(st->uvp, i, j, k) + AT(st->uvp, i - 1, j, k) - AT(st->uvp, i - 1, j + 1, k) + AT(st->uvp, i - 1, j - 1, k) - AT(st->uvp, i - 1, j, k) + AT(st->uwp, i, j, k) - AT(st->uwp, i, j, k + 1) + AT(st->uwp, i, j, k - 1) - AT(st->uwp, i, j, k) + AT(st->uwp, i - 1, j, k) - AT(st->uwp, i - 1, j, k + 1) + AT(st->uwp, i - 1, j, k - 1) - AT(st->uwp, i - 1, j, k)) +
                                AT(st->u2p, i - 1, j, k) - AT(st->u2p, i, j, k));
                    AT(v, i, j, k) =
                        AT(w2, i, j, k) +
                        tauh * (0.25 * (sq(AT(w2, i, j, k) + AT(w2, i, j - 1, k)) - sq(AT(w2, i, j + 1, k) + AT(w2, i, j, k)) +
                                        (AT(w2, i - 1, j, k) + AT(w2, i, j, k)) * (AT(w1, i, j, k) + AT(w1, i, j - 1, k)) -
                                        (AT(w2, i + 1, j, k) + AT(w2, i, j, k)) * (AT(w1, i + 1, j, k) + AT(w1, i + 1, j - 1, k)) +
                                        (AT(w2, i, j, k - 1) + AT(w2, i, j, k)) * (AT(w3, i, j, k) + AT(w3, i, j - 1, k)) -
                                        (AT(w2, i, j, k + 1) + AT(w2, i, j, k)) * (AT(w3, i, j, k + 1) + AT(w3, i, j - 1, k + 1)) +
                                        AT(st->uvp, i - 1, j, k) - AT(st->uvp, i, j, k) + AT(st->uvp, i, j, k) - AT(st->uvp, i + 1, j, k) + AT(st->uvp, i - 1, j - 1, k) - AT(st->uvp, i, j - 1, k) + AT(st->uvp, i, j - 1, k) - AT(st->uvp, i + 1, j - 1, k) + AT(st->vwp, i, j, k - 1) - AT(st->vwp, i, j, k) + AT(st->vwp, i, j, k) - AT(st->vwp, i, j, k + 1) + AT(st->vwp, i, j - 1, k - 1) - AT(st->vwp, i, j - 1, k) + AT(st->vwp, i, j - 1, k) - AT(st->vwp, i, j - 1, k + 1)) +
                                AT(st->v2p, i, j - 1, k) - AT(st->v2p, i, j, k));
                    AT(w, i, j, k) =
                        AT(w3, i, j, k) +
                        tauh * (0.25 * (sq(AT(w3, i, j, k) + AT(w3, i, j, k - 1)) - sq(AT(w3, i, j, k + 1) + AT(w3, i, j, k)) +
                                        (AT(w3, i - 1, j, k) + AT(w3, i, j, k)) * (AT(w1, i, j, k) + AT(w1, i, j, k - 1)) -
                                        (AT(w3, i + 1, j, k) + AT(w3, i, j, k)) * (AT(w1, i + 1, j, k) + AT(w1, i + 1, j, k - 1)) +
                                        (AT(w3, i, j - 1, k) + AT(w3, i, j, k)) * (AT(w2, i, j, k) + AT(w2, i, j, k - 1)) -
                                        (AT(w3, i, j + 1, k) + AT(w3, i, j, k)) * (AT(w2, i, j + 1, k) + AT(w2, i, j + 1, k - 1)) +
                                        AT(st->uwp, i - 1, j, k) - AT(st->uwp, i, j, k) + AT(st->uwp, i, j, k) - AT(st->uwp, i + 1, j, k) + AT(st->uwp, i - 1, j, k - 1) - AT(st->uwp, i, j, k - 1) + AT(st->uwp, i, j, k - 1) - AT(st->uwp, i + 1, j, k - 1) + AT(st->vwp, i, j - 1, k) - AT(st->vwp, i, j, k) + AT(st->vwp, i, j, k) - AT(st->vwp, i, j + 1, k) + AT(st->vwp, i, j - 1, k - 1) - AT(st->vwp, i, j, k - 1) + AT(st->vwp, i, j, k - 1) - AT(st->vwp, i, j + 1, k - 1)) +
                                AT(st->w2p, i, j, k - 1) - AT(st->w2p, i, j, k));
                }
    } else {
        for (int k = 2; k <= nz - 1; k++)
            for (int j = 2; j <= ny - 1; j++)
                for (int i = 2; i <= nx - 1; i++) {
                    AT(u, i, j, k) = AT(w1, i, j, k);
                    AT(v, i, j, k) = AT(w2, i, j, k);
                    AT(w, i, j, k) = AT(w3, i, j, k);
                }
    }
}

static double fmodmax(const double *u, int nx, int ny, int nz)
{
    double m = -1.0;
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++) {
                double r = fabs(AT(u, i, j, k));
                if (r > m) m = r;
            }
    return m;
}

/* solvelpres, main.f:3236-3439 */
void or_solvelpres(or_state *st, int t)
{
    int nx = st->nx, ny = st->ny, nz = st->nz;
    or_level *L = &st->lev[st->ng - 1];
    double *cc = L->cc, *w1 = L->u, *w2 = L->res, *w3 = L->rhs, *a = L->a, *b = L->b, *c = L->c, *ei = L->ei;
    double *u = st->u, *v = st->v, *w = st->w;
    const double tauh2 = (double)st->tauh2, tau = st->tau, h = st->h;
    double ftr = F4(0.0);
    double ftrctr = 0.25 * tauh2 * ftr;
    ftr = 2.0 * ftr * tauh2;
    double r20 = tauh2 * st->mu2;
    double r10 = 0.25 * tauh2 * (st->mu1 - st->mu2);
    double r1 = 2.0 * tauh2 * (st->mu1 - st->mu2);
    double r2 = 2.0 * tauh2 * st->mu2;
    double stdgrav = st->g * (tau * tau) / h;
    double horgrav = st->g1 * (tau * tau) / h;
    double coef = F4(1.0);
    const double *mut = st->mut;
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++) {
                double temp = r2 + r1 * AT(cc, i, j, k) + ftr * coef * AT(mut, i, j, k);
                AT(st->s11, i, j, k) = -AT(st->s11, i, j, k) + temp * (AT(u, i + 1, j, k) - AT(u, i, j, k));
                AT(st->s22, i, j, k) = -AT(st->s22, i, j, k) + temp * (AT(v, i, j + 1, k) - AT(v, i, j, k));
                AT(st->s33, i, j, k) = -AT(st->s33, i, j, k) + temp * (AT(w, i, j, k + 1) - AT(w, i, j, k));
                temp = r20 + r10 * (AT(cc, i, j, k) + AT(cc, i - 1, j, k) + AT(cc, i, j - 1, k) + AT(cc, i - 1, j - 1, k)) +
                       ftrctr * coef * (AT(mut, i, j, k) + AT(mut, i - 1, j, k) + AT(mut, i, j - 1, k) + AT(mut, i - 1, j - 1, k));
                AT(st->s12, i, j, k) = -AT(st->s12, i, j, k) +
                                       (temp) * (AT(u, i, j, k) - AT(u, i, j - 1, k) + AT(v, i, j, k) - AT(v, i - 1, j, k));
                temp = r20 + r10 * (AT(cc, i, j, k) + AT(cc, i - 1, j, k) + AT(cc, i, j, k - 1) + AT(cc, i - 1, j, k - 1)) +
                       ftrctr * coef * (AT(mut, i, j, k) + AT(mut, i - 1, j, k) + AT(mut, i, j, k - 1) + AT(mut, i - 1, j, k - 1));
                AT(st->s13, i, j, k) = -AT(st->s13, i, j, k) +
                                       (temp) * (AT(u, i, j, k) - AT(u, i, j, k - 1) + AT(w, i, j, k) - AT(w, i - 1, j, k));
                temp = r20 + r10 * (AT(cc, i, j, k) + AT(cc, i, j - 1, k) + AT(cc, i, j, k - 1) + AT(cc, i, j - 1, k - 1)) +
                       ftrctr * coef * (AT(mut, i, j, k) + AT(mut, i, j - 1, k) + AT(mut, i, j, k - 1) + AT(mut, i, j - 1, k - 1));
                AT(st->s23, i, j, k) = -AT(st->s23, i, j, k) +
                                       (temp) * (AT(v, i, j, k) - AT(v, i, j, k - 1) + AT(w, i, j, k) - AT(w, i, j - 1, k));
            }
    bc_tensor(st);
    r2 = st->rho2;
    r1 = .5 * (st->rho1 - st->rho2);
    or_abcei(a, b, c, ei, cc, r1, r2, nx, ny, nz);
    const double z0 = 0.0; /* turbst is identically zero on this path */
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++) {
                double turbstx = (z0 + z0) / 2.0, turbsty = (z0 + z0) / 2.0, turbstz = (z0 + z0) / 2.0;
                AT(w1, i, j, k) = AT(u, i, j, k) +
                                  AT(a, i - 1, j, k) * (AT(st->s11, i, j, k) - AT(st->s11, i - 1, j, k) + AT(st->s12, i, j + 1, k) -
                                                        AT(st->s12, i, j, k) + AT(st->s13, i, j, k + 1) - AT(st->s13, i, j, k) - turbstx) +
                                  horgrav;
                AT(w2, i, j, k) = AT(v, i, j, k) +
                                  AT(b, i, j - 1, k) * (AT(st->s12, i + 1, j, k) - AT(st->s12, i, j, k) + AT(st->s22, i, j, k) -
                                                        AT(st->s22, i, j - 1, k) + AT(st->s23, i, j, k + 1) - AT(st->s23, i, j, k) - turbsty) +
                                  stdgrav;
                AT(w3, i, j, k) = AT(w, i, j, k) +
                                  AT(c, i, j, k - 1) * (AT(st->s13, i + 1, j, k) - AT(st->s13, i, j, k) + AT(st->s23, i, j + 1, k) -
                                                        AT(st->s23, i, j, k) + AT(st->s33, i, j, k) - AT(st->s33, i, j, k - 1) - turbstz);
            }
    bc_w(st, w1, w2, w3);
    convective(st, w1, w2, w3);
    bc_v(st);
    const double hi2 = 1.0 / (h * h);
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++)
                AT(w3, i, j, k) = (AT(u, i + 1, j, k) - AT(u, i, j, k) + AT(v, i, j + 1, k) - AT(v, i, j, k) +
                                   AT(w, i, j, k + 1) - AT(w, i, j, k)) *
                                  hi2;
    if (st->mg_mode == 0) {
        or_mglin(st);
    } else {
        or_mg_solve_tol(st);
    }
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++) {
                AT(u, i, j, k) = AT(u, i, j, k) - (AT(w1, i, j, k) - AT(w1, i - 1, j, k)) * AT(a, i - 1, j, k);
                AT(v, i, j, k) = AT(v, i, j, k) - (AT(w1, i, j, k) - AT(w1, i, j - 1, k)) * AT(b, i, j - 1, k);
                AT(w, i, j, k) = AT(w, i, j, k) - (AT(w1, i, j, k) - AT(w1, i, j, k - 1)) * AT(c, i, j, k - 1);
            }
    bc_v(st);
    if (t <= 1 || t % 10 == 0) {
        for (int k = 2; k <= nz - 1; k++)
            for (int j = 2; j <= ny - 1; j++)
                for (int i = 2; i <= nx - 1; i++)
                    AT(w3, i, j, k) = (AT(u, i + 1, j, k) - AT(u, i, j, k) + AT(v, i, j + 1, k) - AT(v, i, j, k) +
                                       AT(w, i, j, k + 1) - AT(w, i, j, k)) *
                                      hi2;
        double epsdiv = h * fmodmax(w3, nx, ny, nz);
        if (epsdiv > 25.0 && st->ntop < 30) {
            st->ntop += 2; st->npre += 2; st->npost += 2;
        } else if (epsdiv < 2.5 && st->ntop > 4) {
            st->npre -= 1; st->ntop -= 1; st->npost -= 1;
        }
    }
}

/* one pass of the main loop body, main.f:21-71 (printout is output only) */
void or_step(or_state *st, int t)
{
    or_vofmethod(st, t);
    if (st->hswitch == 1) or_calch3d(st);  /* main.f:30 */
    if (st->iturb == 1 || st->iturb == 2) or_eddyvisc(st);  /* main.f:35 */
    if (st->iturb == 3 || st->iturb == 4) or_kepsmodel(st, t);  /* main.f:39-40 */
    or_checktime(st, t);
    if (st->surfacet == 2) or_csf(st);
    else if (st->surfacet == 5) or_sgip(st);
    else if (st->surfacet == 6) or_sgip2(st);
    or_solvelpres(st, t);
}

/* ------------------------------------------------------------------ */
/* set-up: readinput main.f:185-272, memoryaloc main.f:274-322,
 * nondim main.f:341-446, smooth2 main.f:5678-5752, initialize main.f:121-183 */

static int read_first_token(FILE *f, char *tok, int n)
{
    char line[512];
    if (!fgets(line, sizeof line, f)) return -1;
    char *p = line;
    while (*p == ' ' || *p == '\t') p++;
    int m = 0;
    while (*p && *p != ' ' && *p != '\t' && *p != '\n' && *p != '\r' && *p != ',' && m < n - 1) tok[m++] = *p++;
    tok[m] = 0;
    return m > 0 ? 0 : -1;
}

static void smooth2(or_state *st)
{
    int nx = st->nx, ny = st->ny, nz = st->nz;
    double *cc = st->lev[st->ng - 1].cc;
    const double h = st->h;
    double x0 = (double)(nx - 2) * h / 4.0;
    if (st->z0_si != 1000.0) st->z0 = st->z0_si / st->length_si;
    if (st->y0_si != 1000.0) st->y0 = st->y0_si / st->length_si;
    st->y0 = st->y0 + h / 2.0;
    const double y0 = st->y0, z0 = st->z0, r2 = st->radius * st->radius;
    /* xx(1), yy(1), zz(1) are never assigned by the reference (main.f:5683);
     * the natural extension (1-1.5)h is used. */
    double *xx = malloc(sizeof(double) * (nx + 1));
    double *yy = malloc(sizeof(double) * (ny + 1));
    double *zz = malloc(sizeof(double) * (nz + 1));
    for (int i = 1; i <= nx; i++) xx[i] = ((double)i - 1.5) * h;
    for (int j = 1; j <= ny; j++) yy[j] = ((double)j - 1.5) * h;
    for (int k = 1; k <= nz; k++) zz[k] = ((double)k - 1.5) * h;
    const int index = 30;
    const double inc = 1.0 / (double)(index * index * index);
    for (int i = 2; i <= nx - 1; i++)
        for (int j = 2; j <= ny - 1; j++)
            for (int k = 2; k <= nz - 1; k++) {
                double dmin = dmin1(sq(xx[i - 1] - x0), sq(xx[i] - x0)) + dmin1(sq(yy[j - 1] - y0), sq(yy[j] - y0)) +
                              dmin1(sq(zz[k - 1] - z0), sq(zz[k] - z0));
                if (!(dmin < r2)) {
                    AT(cc, i, j, k) = 0.0;
                    continue;
                }
                double dmax = dmax1(sq(xx[i - 1] - x0), sq(xx[i] - x0)) + dmax1(sq(yy[j - 1] - y0), sq(yy[j] - y0)) +
                              dmax1(sq(zz[k - 1] - z0), sq(zz[k] - z0));
                if (!(dmax > r2)) {
                    AT(cc, i, j, k) = 1.0;
                    continue;
                }
                double s = 0.0;
                for (int l = 1; l <= index; l++)
                    for (int m = 1; m <= index; m++)
                        for (int n = 1; n <= index; n++) {
                            double px = xx[i - 1] + ((double)l - .5) / (double)index * h - x0;
                            double py = yy[j - 1] + ((double)m - .5) / (double)index * h - y0;
                            double pz = zz[k - 1] + ((double)n - .5) / (double)index * h - z0;
                            if (sq(px) + sq(py) + sq(pz) < r2) s = s + inc;
                        }
                AT(cc, i, j, k) = s;
            }
    free(xx); free(yy); free(zz);
    bc_c(st, cc);
    size_t n = (size_t)nx * ny * nz;
    for (size_t p = 0; p < n; p++) {
        st->u[p] = 0.0;
        st->v[p] = st->yvel1 * cc[p] + st->yvel2 * (1.0 - cc[p]);
        st->w[p] = st->zvel1 * cc[p] + st->zvel2 * (1.0 - cc[p]);
    }
}

static void nondim(or_state *st)
{
    size_t n = (size_t)st->nx * st->ny * st->nz;
    for (size_t p = 0; p < n; p++) st->cmu[p] = 0.09;
    st->lengthscale = st->length_si;
    double vs = sqrt((st->xvel1_si - st->xvel2_si) * (st->xvel1_si - st->xvel2_si) +
                     (st->yvel1_si - st->yvel2_si) * (st->yvel1_si - st->yvel2_si) +
                     (st->zvel1_si - st->zvel2_si) * (st->zvel1_si - st->zvel2_si));
    double c_drag = 1.0;
    double vel2 = sqrt(st->xvel2_si * st->xvel2_si + st->yvel2_si * st->yvel2_si + st->zvel2_si * st->zvel2_si);
    double gtot = sqrt(st->grav_si * st->grav_si + st->grav1_si * st->grav1_si);
    if (vs == 0.0 && gtot != 0.0) {
        vs = sqrt((fabs(gtot) * st->lengthscale) * fabs(st->rho1_si - st->rho2_si) / st->rho2_si * (1.0 / c_drag));
        double re2 = vel2 * st->lengthscale * st->rho2_si / st->visc2_si;
        if (re2 <= 10.0 && vs == 0.0) c_drag = 10.0 / re2;
        vs = sqrt((fabs(gtot) * st->lengthscale) * fabs(st->rho1_si - st->rho2_si) / st->rho2_si * (1.0 / c_drag));
    }
    if (vs == 0.0) vs = 1.0;
    st->vscale = vs;
    st->timescale = st->lengthscale / vs;
    st->lengthsize = st->length_si / st->lengthscale;
    st->h = st->lengthsize / (double)(st->nx - 2);
    st->time = 0.0;
    st->tau_si = st->maxcfl * F4(0.8) * (st->length_si / (double)(st->nx - 2)) / vs;
    st->tau = st->tau_si / st->timescale;
    st->tauh = st->tau / st->h;
    st->tauh2 = (int)(st->tau / (st->h * st->h));
    st->tswap = 0;
    st->radius = st->radius_si / st->lengthscale;
    st->xvel1 = st->xvel1_si / vs; st->yvel1 = st->yvel1_si / vs; st->zvel1 = st->zvel1_si / vs;
    st->xvel2 = st->xvel2_si / vs; st->yvel2 = st->yvel2_si / vs; st->zvel2 = st->zvel2_si / vs;
    st->y0 = (double)(st->ny - 2) * st->h / 2.0;
    st->z0 = (double)(st->nz - 2) * st->h / 2.0;
    st->rho1 = 1.0;
    st->rho2 = st->rho2_si / st->rho1_si;
    double denom = st->rho1_si * (st->lengthscale * st->lengthscale) / st->timescale;
    st->mu1 = st->visc1_si / denom;
    st->mu2 = st->visc2_si / denom;
    denom = st->rho1_si * (st->lengthscale * st->lengthscale * st->lengthscale) / (st->timescale * st->timescale);
    st->sigma = st->sigma_si / denom;
    denom = st->lengthscale / (st->timescale * st->timescale);
    st->g = st->grav_si / denom;
    st->g1 = st->grav1_si / denom;
    memset(st->mut, 0, sizeof(double) * n);
}

static double *zalloc(size_t n)
{
    double *p = (double *)calloc(n, sizeof(double));
    return p;
}

or_state *or_create(const char *path, char *errbuf, int errlen)
{
    FILE *f = fopen(path, "r");
    if (!f) {
        if (errbuf) snprintf(errbuf, errlen, "cannot open %s", path);
        return NULL;
    }
    or_state *st = (or_state *)calloc(1, sizeof(or_state));
    char tok[128];
    double vals[38];
    for (int q = 0; q < 38; q++) {
        if (read_first_token(f, tok, sizeof tok) != 0) {
            if (errbuf) snprintf(errbuf, errlen, "inputdata: missing value on line %d", q + 1);
            fclose(f); free(st);
            return NULL;
        }
        vals[q] = strtod(tok, NULL);
    }
    fclose(f);
    int q = 0;
    st->nx = (int)vals[q++]; st->ny = (int)vals[q++]; st->nz = (int)vals[q++];
    st->tmax = (int)vals[q++]; st->nfilter = (int)vals[q++]; st->cycles = (int)vals[q++];
    st->omega = vals[q++];
    st->npre = (int)vals[q++]; st->ntop = (int)vals[q++]; st->npost = (int)vals[q++];
    st->restart = (int)vals[q++];
    st->maxcfl = vals[q++];
    st->iturb = (int)vals[q++]; st->convectype = (int)vals[q++]; st->hswitch = (int)vals[q++];
    st->xvel1_si = vals[q++]; st->yvel1_si = vals[q++]; st->zvel1_si = vals[q++];
    st->xvel2_si = vals[q++]; st->yvel2_si = vals[q++]; st->zvel2_si = vals[q++];
    st->rho2_si = vals[q++]; st->rho1_si = vals[q++]; st->sigma_si = vals[q++];
    st->visc2_si = vals[q++]; st->visc1_si = vals[q++]; st->grav_si = vals[q++]; st->grav1_si = vals[q++];
    st->icase = (int)vals[q++]; st->ibc = (int)vals[q++];
    st->length_si = vals[q++]; st->radius_si = vals[q++]; st->y0_si = vals[q++]; st->z0_si = vals[q++];
    st->tprint = (int)vals[q++]; st->ntecplot = (int)vals[q++]; st->tsave = (int)vals[q++]; st->surfacet = (int)vals[q++];
    if (st->icase != 2) {
        if (errbuf) snprintf(errbuf, errlen, "icase=%d: only the droplet case (icase=2) is on this path", st->icase);
        free(st);
        return NULL;
    }
    st->idrop = 1;
    /* memoryaloc: levels n -> n/2+1 while the smallest extent exceeds 4.
     * The reference STOPs unless nn = 2**ng + 2; that check moves to mglin
     * (or_mglin), its only user, so that m*2**k + 2 grids serve the
     * tolerance-mode solver: their halving stops where an interior extent
     * turns odd. */
    int nn = st->nx < st->ny ? st->nx : st->ny;
    if (st->nz < nn) nn = st->nz;
    int ng = 0;
    {
        int lx = st->nx, ly = st->ny, lz = st->nz;
        for (;;) {
            if (ng >= OR_MAXLEV) {
                if (errbuf) snprintf(errbuf, errlen, "too many multigrid levels (%d)", nn);
                free(st);
                return NULL;
            }
            ng++;
            int m = lx < ly ? lx : ly;
            if (lz < m) m = lz;
            if (m <= 4 || ((lx - 2) & 1) || ((ly - 2) & 1) || ((lz - 2) & 1)) break;
            lx = lx / 2 + 1; ly = ly / 2 + 1; lz = lz / 2 + 1;
        }
        int j = nn - 2, ngr = 1;
        while (j > 2) { j = j / 2; ngr++; }
        st->ref_grid = (1 << ngr) + 2 == nn && ngr == ng;
    }
    st->ng = ng;
    int lx = st->nx, ly = st->ny, lz = st->nz;
    for (int g = ng; g >= 1; g--) {
        or_level *L = &st->lev[g - 1];
        L->nx = lx; L->ny = ly; L->nz = lz;
        size_t n = (size_t)lx * ly * lz;
        L->rhs = zalloc(n); L->res = zalloc(n); L->u = zalloc(n); L->cc = zalloc(n);
        L->a = zalloc(n); L->b = zalloc(n); L->c = zalloc(n); L->ei = zalloc(n);
        lx = lx / 2 + 1; ly = ly / 2 + 1; lz = lz / 2 + 1;
    }
    size_t n = (size_t)st->nx * st->ny * st->nz;
    st->u = zalloc(n); st->v = zalloc(n); st->w = zalloc(n);
    st->s11 = zalloc(n); st->s22 = zalloc(n); st->s33 = zalloc(n);
    st->s12 = zalloc(n); st->s13 = zalloc(n); st->s23 = zalloc(n);
    st->rhon = zalloc(n); st->mun = zalloc(n); st->mut = zalloc(n);
    st->ktur = zalloc(n); st->eps = zalloc(n); st->cmu = zalloc(n);
    st->p_prev = zalloc(n);
    st->hxmin = zalloc(n); st->hxmax = zalloc(n); st->hymin = zalloc(n);
    st->hymax = zalloc(n); st->hzmin = zalloc(n); st->hzmax = zalloc(n);
    st->u2p = zalloc(n); st->v2p = zalloc(n); st->w2p = zalloc(n);
    st->uvp = zalloc(n); st->uwp = zalloc(n); st->vwp = zalloc(n);
    st->lmix = zalloc((size_t)st->nx);
    st->mg_mode = 0;
    st->mg_nu1 = 2; st->mg_nu2 = 2; st->mg_maxcycles = 200; st->mg_tol = 1e-8; st->mg_guess = 0;
    nondim(st);
    smooth2(st);
    bc_c(st, st->lev[ng - 1].cc);
    bc_v(st);
    or_calcrho(st);
    return st;
}

void or_destroy(or_state *st)
{
    if (!st) return;
    for (int g = 0; g < st->ng; g++) {
        or_level *L = &st->lev[g];
        free(L->rhs); free(L->res); free(L->u); free(L->cc); free(L->a); free(L->b); free(L->c); free(L->ei);
    }
    free(st->u); free(st->v); free(st->w);
    free(st->s11); free(st->s22); free(st->s33); free(st->s12); free(st->s13); free(st->s23);
    free(st->rhon); free(st->mun); free(st->mut); free(st->ktur); free(st->eps); free(st->cmu);
    free(st->p_prev);
    free(st->hxmin); free(st->hxmax); free(st->hymin); free(st->hymax); free(st->hzmin); free(st->hzmax);
    free(st->u2p); free(st->v2p); free(st->w2p); free(st->uvp); free(st->uwp); free(st->vwp); free(st->lmix);
    free(st);
}

double *or_field(or_state *st, int id)
{
    or_level *L = &st->lev[st->ng - 1];
    switch (id) {
    case OR_F_CC: return L->cc;
    case OR_F_U: return st->u;
    case OR_F_V: return st->v;
    case OR_F_W: return st->w;
    case OR_F_P: return L->u;
    case OR_F_S11: return st->s11;
    case OR_F_S22: return st->s22;
    case OR_F_S33: return st->s33;
    case OR_F_S12: return st->s12;
    case OR_F_S13: return st->s13;
    case OR_F_S23: return st->s23;
    case OR_F_RHON: return st->rhon;
    case OR_F_MUN: return st->mun;
    case OR_F_MUT: return st->mut;
    case OR_F_KTUR: return st->ktur;
    case OR_F_EPS: return st->eps;
    case OR_F_A: return L->a;
    case OR_F_B: return L->b;
    case OR_F_C: return L->c;
    case OR_F_EI: return L->ei;
    case OR_F_RHS: return L->rhs;
    case OR_F_RES: return L->res;
    case OR_F_PPREV: return st->p_prev;
    case OR_F_HXMIN: return st->hxmin;
    case OR_F_HXMAX: return st->hxmax;
    case OR_F_HYMIN: return st->hymin;
    case OR_F_HYMAX: return st->hymax;
    case OR_F_HZMIN: return st->hzmin;
    case OR_F_HZMAX: return st->hzmax;
    case OR_F_U2P: return st->u2p;
    case OR_F_V2P: return st->v2p;
    case OR_F_W2P: return st->w2p;
    case OR_F_UVP: return st->uvp;
    case OR_F_UWP: return st->uwp;
    case OR_F_VWP: return st->vwp;
    case OR_F_CMU: return st->cmu;
    default: return NULL;
    }
}

double *or_level_array(or_state *st, int level, int which)
{
    if (level < 1 || level > st->ng) return NULL;
    or_level *L = &st->lev[level - 1];
    double *arr[8] = {L->rhs, L->res, L->u, L->cc, L->a, L->b, L->c, L->ei};
    if (which < 0 || which > 7) return NULL;
    return arr[which];
}

int or_level_dims(or_state *st, int level, int *dims3)
{
    if (level < 1 || level > st->ng) return -1;
    or_level *L = &st->lev[level - 1];
    dims3[0] = L->nx; dims3[1] = L->ny; dims3[2] = L->nz;
    return 0;
}
