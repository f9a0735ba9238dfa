/*
 * droplet_oracle.h -- CPU restatement of the droplet solver's per-time-step
 * path (farzan-t/3D_Droplet_simulation_2007, main.f).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the GPU product
 * (libdroplet_gpu.so) never links, loads or calls it.  Only tests/, the
 * smoke() entry point and bench.py's cpu_baseline leg may use it.
 *
 * Every routine states the main.f line range it restates.  Arrays follow the
 * reference layout: FP64, column-major (nx,ny,nz) with one ghost layer, the
 * x index unit-stride, Fortran 1-based indices in the code.
 */
#ifndef DROPLET_ORACLE_H
#define DROPLET_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXLEV 16

typedef struct or_level {
    int nx, ny, nz;
    double *rhs, *res, *u, *cc, *a, *b, *c, *ei;   /* memoryaloc, main.f:305-318 */
} or_level;

typedef struct or_state {
    /* /DIMENSION/ (input.h:29-30) */
    int nx, ny, nz;
    /* inputdata.in, in readinput order (main.f:188-234) */
    int tmax, nfilter, cycles;
    double omega;
    int npre, ntop, npost, restart;
    double maxcfl;
    int iturb, convectype, hswitch;
    double xvel1_si, yvel1_si, zvel1_si, xvel2_si, yvel2_si, zvel2_si;
    double rho2_si, rho1_si, sigma_si, visc2_si, visc1_si, grav_si, grav1_si;
    int icase, ibc;
    double length_si, radius_si, y0_si, z0_si;
    int tprint, ntecplot, tsave, surfacet;
    /* derived switches (main.f:243-268) */
    int ijet, idrop, ibubble, isheet;
    /* multigrid heap (memoryaloc, main.f:274-322); lev[0] = Fortran level 1 */
    int ng;
    or_level lev[OR_MAXLEV];
    /* /DINEM/ */
    double tau_si, timescale, lengthscale, lengthsize, vscale;
    /* /PHYSICS/ scalars */
    double tau, time, radius, h, rho1, rho2, mu1, mu2, sigma, g, g1;
    double y0, z0, xvel1, yvel1, zvel1, xvel2, yvel2, zvel2;
    /* /DIFF/: tauh2 is declared INTEGER in input.h, so every assignment
     * tauh2 = tau/(h*h) truncates toward zero (main.f:389,2164,2171) */
    double tauh;
    int tauh2;
    int tswap;
    /* /SURFCOM/ */
    double *u, *v, *w, *s11, *s22, *s33, *s12, *s13, *s23;
    /* /PHYSICS/ arrays and /TURBULENCE/ arrays used on the path */
    double *rhon, *mun, *mut, *ktur, *eps, *cmu;
    /* pressure solver selection: 0 = reference FMG schedule (mglin),
     * 1 = FP64 V-cycles to a relative-residual tolerance,
     * 2 = FP32 V-cycles inside an FP64 defect-correction loop */
    int mg_mode;
    int mg_nu1, mg_nu2, mg_maxcycles, mg_guess;   /* guess: 0 zero, 1 previous p */
    double mg_tol;
    int mg_cycles_last;
    double mg_relres_last, mg_bnorm_last;
    double *p_prev;
    /* last step diagnostics */
    double cfl_last;
    /* tolerance-mode options: outer 0 = defect correction, 1 = (flexible)
     * preconditioned CG; prolong 0 = injection + 8-average restriction,
     * 1 = trilinear + its transpose */
    int mg_outer, mg_prolong;
    /* 1: a partly filled cell whose Youngs normal vanishes (no filled
     * neighbour, or a symmetric neighbourhood) is advected with a plane
     * normal to the sweep direction instead of AL3D/FL3D's 0/0 (NaN) of
     * swp main.f:848-1000.  0 (default): the reference's arithmetic. */
    int vof_guard;
    /* tolerance-mode smoother: SOR weight of the red-black passes (1 =
     * Gauss-Seidel) and the sweep count (pre and post) on the levels below
     * the finest (0: mg_nu1 / mg_nu2 there too) */
    double mg_omega;
    int mg_nuc;
    /* 1: the grid has the reference's form nn = 2**ng + 2 (memoryaloc's
     * check, main.f:274-322), so the FMG schedule (mglin) may run; 0: an
     * m*2**k + 2 grid, tolerance mode only */
    int ref_grid;
    /* /NEWW/ interface heights of CalcH3d (main.f:1312-1810), read by SGIP
     * and SGIP2; allocated zeroed, only the interior is ever written */
    double *hxmin, *hxmax, *hymin, *hymax, *hzmin, *hzmax;
    /* a STOP of the reference on the last step: 0 none, 1 CalcH3d's "NO
     * MATCH CASES FOUND", 2 its "PROBLEM finding MIN and MAX unit vector
     * components" (main.f:1582-1583, 1641-1644) */
    int stop_code;
    /* /TURBULENCE/ Reynolds stresses of the anisotropic k-epsilon model
     * (uv2p, main.f:2032-2155; zero unless iturb = 4, nondim main.f:436-441)
     * and eddyvisc's mixing length lmix(NMX) (main.f:1828-1885) */
    double *u2p, *v2p, *w2p, *uvp, *uwp, *vwp;
    double *lmix;
    /* tolerance-mode smoother on the levels strictly between the finest and
     * the coarsest: mg_blk[0..2] > 0 smooths each tile of that many cells
     * (x, y, z) on its own: the tile grown by mg_blk_ov cells is smoothed
     * with the cells around it held at their values from before the sweeps,
     * and the tile's own cells are kept (restricted additive Schwarz;
     * ov = 0: block Jacobi), red-black SOR inside; only on levels of more
     * than mg_blk_min and (mg_blk_max > 0) at most mg_blk_max interior
     * cells.  0 (default): global red-black passes. */
    int mg_blk[3];
    int mg_blk_min;
    int mg_blk_ov;
    int mg_blk_max;
    int mg_blk_nu;  /* sweeps (pre and post) on the tile levels; 0: the levels' own count */
    int mg_blk2[3]; /* the tile of the tile levels of more than mg_blk2_min cells (0: mg_blk) */
    int mg_blk2_min;
} or_state;

/* field ids shared with the GPU library's dg_field ids */
enum {
    OR_F_CC = 0, OR_F_U, OR_F_V, OR_F_W, OR_F_P,
    OR_F_S11, OR_F_S22, OR_F_S33, OR_F_S12, OR_F_S13, OR_F_S23,
    OR_F_RHON, OR_F_MUN, OR_F_MUT, OR_F_KTUR, OR_F_EPS,
    OR_F_A, OR_F_B, OR_F_C, OR_F_EI, OR_F_RHS, OR_F_RES,
    OR_F_PPREV,
    OR_F_HXMIN, OR_F_HXMAX, OR_F_HYMIN, OR_F_HYMAX, OR_F_HZMIN, OR_F_HZMAX,
    OR_F_U2P, OR_F_V2P, OR_F_W2P, OR_F_UVP, OR_F_UWP, OR_F_VWP,
    OR_F_CMU,
    OR_NFIELDS
};

/* ---- set-up: readinput + memoryaloc + nondim + IC (main.f:121-322,341-446,5678-5752) */
or_state *or_create(const char *inputdata_path, char *errbuf, int errlen);
void      or_destroy(or_state *st);
double   *or_field(or_state *st, int id);
double   *or_level_array(or_state *st, int level, int which);  /* which: 0 rhs,1 res,2 u,3 cc,4 a,5 b,6 c,7 ei */
int       or_level_dims(or_state *st, int level, int *dims3);

/* ---- one time step and its parts (main.f:19-73) */
void or_step(or_state *st, int t);
void or_vofmethod(or_state *st, int t);           /* main.f:795-824 */
void or_kepsmodel(or_state *st, int t);           /* main.f:1888-2026 (iturb 3, 4) */
void or_eddyvisc(or_state *st);                   /* main.f:1828-1885 (iturb 1, 2) */
void or_uv2p(or_state *st);                       /* main.f:2032-2155 (iturb 4) */
void or_checktime(or_state *st, int t);           /* main.f:2156-2175 */
void or_csf(or_state *st);                        /* main.f:2363-2481 */
void or_calch3d(or_state *st);                    /* main.f:1312-1810 */
void or_sgip(or_state *st);                       /* main.f:2835-3027 */
void or_sgip2(or_state *st);                      /* main.f:3029-3166 */
void or_solvelpres(or_state *st, int t);          /* main.f:3236-3439 */
void or_calcrho(or_state *st);                    /* main.f:2177-2190 */
double or_cfl(or_state *st);                      /* main.f:4848-4881 */

/* ---- pure kernels with the reference's argument lists */
double or_al3d(double b1, double b2, double b3, double cc);                         /* main.f:4562 */
double or_fl3d(double m1, double m2, double m3, double alpha, double r0, double dr0); /* main.f:4639 */
void or_relax(double *u, double lengthsize, const double *rhs, const double *a, const double *b,
              const double *c, const double *ei, int nx, int ny, int nz, double omega);      /* main.f:3956 */
void or_resid(double lengthsize, double *res, const double *u, const double *rhs, const double *a,
              const double *b, const double *c, int nx, int ny, int nz);                      /* main.f:3902 */
void or_rstrct(double *uc, const double *uf, int ncx, int ncy, int ncz);                   /* main.f:3765 */
void or_interp(double *uf, const double *uc, int nfx, int nfy, int nfz);                   /* main.f:3866 */
void or_addint(double *uf, const double *uc, double *res, int nfx, int nfy, int nfz);      /* main.f:3934 */
void or_acfromac(double *a, double *b, double *c, double *ei, const double *a2, const double *b2,
                 const double *c2, int nx, int ny, int nz);                                /* main.f:3793 */
void or_slvsml(double *u, const double *rhs, const double *a, const double *b, const double *c,
               const double *ei, int nx, int ny, int nz, int ntop, double omega, double lengthsize); /* main.f:3846 */
void or_abcei(double *a, double *b, double *c, double *ei, const double *cc, double r1, double r2,
              int nx, int ny, int nz);                                                      /* main.f:3582 */
void or_bc_scalar(double *s, int nx, int ny, int nz);                                      /* main.f:4298 */
void or_filterc27(const double *cc, double *cf, int nx, int ny, int nz);                   /* main.f:5473 */
void or_mglin(or_state *st);                                                               /* main.f:3645 */

/* ---- tolerance-mode pressure solve (no reference counterpart): solves the
 * finest-level system held in lev[ng-1] (rhs, a, b, c, ei) into lev[ng-1].u */
int  or_mg_solve_tol(or_state *st);
/* the tile smoother of the levels between the finest and the coarsest (0, 0, 0: global passes) */
void or_set_block_smoother(or_state *st, int bx, int by, int bz, int min_cells);
void or_set_block_overlap(or_state *st, int ov);
void or_set_block_max(or_state *st, int max_cells);
void or_set_block_sweeps(or_state *st, int nu);
void or_set_block_large(or_state *st, int bx, int by, int bz, int min_cells);
double or_resid_norm(or_state *st, double *bnorm_out);

#ifdef __cplusplus
}
#endif
#endif
