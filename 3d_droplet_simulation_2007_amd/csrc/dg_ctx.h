// dg_ctx.h -- device-resident solver state and the operator entry points
// shared between the kernel files and the C-ABI layer.
#pragma once

#include "dg_internal.h"

namespace dg {

constexpr int MAXLEV = 16;

// One multigrid level.  The eight FP64 arrays mirror the reference's heap
// slots (memoryaloc, main.f:305-318), including their reuse as scratch by
// VOF (vof1/2/3 = ei/res/rhs, main.f:806-816), csf (w11/w22/w33 =
// u/res/rhs, main.f:52) and solvelpres (w1/w2/w3 = u/res/rhs, main.f:68).
struct Level {
    Grid g;
    double *rhs = nullptr, *res = nullptr, *u = nullptr, *cc = nullptr;
    double *a = nullptr, *b = nullptr, *c = nullptr, *ei = nullptr;
    // tolerance-mode work arrays
    float *a32 = nullptr, *b32 = nullptr, *c32 = nullptr;
    float *u32 = nullptr, *f32 = nullptr, *r32 = nullptr;
    double *u64 = nullptr, *f64 = nullptr, *r64 = nullptr;
    // red-black split storage of the V-cycle (levels above the coarsest):
    // u and f per colour (FP32 with the FP32 V-cycle, FP64 with the FP64
    // one), a/b/c per colour in FP16, FP32 or FP64
    float *rb_u = nullptr, *rb_f = nullptr;
    float *rb_w = nullptr;  // the tile smoother's post-smoothing result (out of place)
    double *rb_u64 = nullptr, *rb_f64 = nullptr, *rb_w64 = nullptr;
    void *rb_coef = nullptr;
    int rb_ct = -1;  // coefficient type rb_coef was allocated for: RB_CT_F32, RB_CT_F16, RB_CT_F64 (-1: none)
};

// coefficient types of the split V-cycle
enum { RB_CT_F32 = 0, RB_CT_F16 = 1, RB_CT_F64 = 2 };

struct Scalars {
    // /PHYSICS/, /DIFF/, /DINEM/ values used on the path
    double tau, time, h, tauh;
    int tauh2;  // INTEGER in input.h: tau/(h*h) truncated
    int tswap;
    double rho1, rho2, mu1, mu2, sigma, g, g1;
    double xvel1, yvel1, zvel1, xvel2, yvel2, zvel2, radius, y0, z0;
    double lengthscale, lengthsize, timescale, vscale, tau_si;
    double maxcfl, omega;
    int cycles, npre, ntop, npost, iturb, convectype, surfacet, hswitch, tprint;
};

struct Ctx {
    Grid g{};
    int ng = 0;
    Level lev[MAXLEV];
    double *u = nullptr, *v = nullptr, *w = nullptr;
    double *s11 = nullptr, *s22 = nullptr, *s33 = nullptr, *s12 = nullptr, *s13 = nullptr, *s23 = nullptr;
    double *rhon = nullptr, *mun = nullptr, *mut = nullptr, *ktur = nullptr, *eps = nullptr, *cmu = nullptr;
    double *ccf = nullptr;     // csf's filtered volume fraction (automatic array in csf)
    // CalcH3d's interface heights (/NEWW/, input.h:90-95); allocated with
    // hswitch = 1 or SGIP/SGIP2 surface tension, only the interior written
    double *hxmin = nullptr, *hxmax = nullptr, *hymin = nullptr, *hymax = nullptr, *hzmin = nullptr, *hzmax = nullptr;
    // uv2p's Reynolds stresses (/TURBULENCE/, input.h:36-43); allocated
    // with iturb = 4, zero otherwise in the reference (nondim main.f:436-441)
    double *u2p = nullptr, *v2p = nullptr, *w2p = nullptr, *uvp = nullptr, *uwp = nullptr, *vwp = nullptr;
    double *lmix = nullptr;    // eddyvisc's mixing length lmix(NMX): nx doubles, then nx of scratch
    double *pprev = nullptr;   // pressure of the previous step (tolerance-mode guess)
    double *pcg_r = nullptr, *pcg_p = nullptr, *pcg_q = nullptr;  // pcg_p: the ring of search directions
    double *red_partial = nullptr;  // RED_BLOCKS doubles
    double *dscal = nullptr;        // small device scalar scratch (64 doubles)
    double *hscal = nullptr;        // pinned host mirror of dscal
    double *hpoll = nullptr;        // mapped coherent host slots: [1] the count of PCG iterations done, [2 + 2 (n & 1)] and the next: r.r and the stopping flag of iteration n
    double *dpoll = nullptr;        // the same slots' device address
    double *coarse_M = nullptr;     // Cholesky factor of the coarsest level's -A (tolerance mode), n0 x n0
    size_t coarse_msize = 0;        // doubles allocated for coarse_M
    Scalars s{};
    // pressure-solver options
    int mg_mode = 0;       // 0 reference FMG schedule, 1 FP64 PCG, 2 FP32-preconditioned FP64 PCG
    int mg_nu1 = 2, mg_nu2 = 2, mg_maxit = 200, mg_guess = 0;
    double mg_tol = 1e-8;
    double mg_omega = 1.0;  // SOR weight of the tolerance-mode smoother (1: Gauss-Seidel)
    int mg_nuc = 0;         // sweeps (pre and post) below the finest level; 0: mg_nu1 / mg_nu2
    int ordered_sums = 1;  // reproduce the reference's sequential outflow sums
    int vof_guard = 0;     // 1: plane normal to the sweep where the Youngs normal vanishes
    int use_graphs = 1;      // replay each PCG iteration as a captured HIP graph
    double pcg_lookahead = 1e3;  // PCG iteration n+1 is queued before n's residual reaches the host while
                                 // the last residual the host has seen exceeds this multiple of the
                                 // tolerance (<= 0: never)
    int ref_grid = 1;        // the grid has the reference's form nn = 2**ng + 2 (mglin's schedule may run)
    int rb_split = 1;        // red-black split storage of the V-cycle levels (modes 1 and 2)
    int rb_tail = 2;         // split levels of <= 4096 cells: 0 per-level launches, 1 one-workgroup tail, 2 LDS tail
    int rb_tiles = 0;        // split levels between the finest and the tail smoothed as tiles (one launch per
                             // level visit, restricted additive Schwarz with red-black SOR in each tile):
                             // 0 none, 1 those of at most 2^18 cells, 2 all
    int rb_tile_nu = 0;      // sweeps (pre and post) on the tile levels; 0: the levels' own count
    int coef16_ok = 0;       // face coefficients fit FP16 (set from the fluids' densities)
    int top_coef_ready = 0;  // solvelpres's abcei wrote the finest level's split coefficients for this solve
    int coarse_coef_ready = 0;  // acfromac wrote the coarse levels' split coefficients for this solve
    hipGraphExec_t pcg_graph[2] = {nullptr, nullptr};
    int pcg_graph_key[13] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    // diagnostics of the last step
    int mg_iters_last = 0;
    double mg_relres_last = 0.0, cfl_last = 0.0, epsdiv_last = 0.0;
    hipStream_t stream = nullptr;
};

Ctx *ctx();  // the process-wide context (the reference keeps one COMMON state)

// ---- operators (return 0 on success, -1 on error with set_error) ----
int op_bc_c(Ctx &c, double *cc);
int op_bc_v(Ctx &c);
int op_bc_w(Ctx &c, double *w1, double *w2, double *w3);
int op_bc_tensor(Ctx &c);
int op_bc_scalar(const Grid &g, hipStream_t s, double *a);
int op_bc_out(const Grid &g, hipStream_t s, double *a);
int op_bc_resid(const Grid &g, hipStream_t s, double *a);
int op_calcrho(Ctx &c);
int op_vofmethod(Ctx &c, int t, bool step = false);
int op_kepsmodel(Ctx &c, int t);
int op_cfl(Ctx &c, double *out_host);
int op_checktime(Ctx &c, int t);
int op_csf(Ctx &c, bool w_interior = true);
int op_calch3d(Ctx &c);                           // CalcH3d main.f:1312-1810
int op_eddyvisc(Ctx &c);                          // eddyvisc main.f:1828-1885 (iturb 1, 2)
int op_uv2p(Ctx &c);                              // uv2p main.f:2032-2155 (iturb 4)
// smooth2 (main.f:5678-5752) for n droplets: centres, radii (nondimensional)
// and liquid velocities in device-visible host order (x0, y0, z0, r, u, v, w)
int op_droplets(Ctx &c, int n, const double *params);
int op_sgip(Ctx &c, bool w_interior = true);      // SGIP main.f:2835-3027
int op_sgip2(Ctx &c, bool w_interior = true);     // SGIP2 main.f:3029-3166
int op_solvelpres(Ctx &c, int t);
int op_smooth2(Ctx &c);

// multigrid, reference schedule (main.f:3645-3997)
int mg_relax(const Grid &g, hipStream_t s, double *u, double lengthsize, const double *rhs, const double *a,
             const double *b, const double *c, const double *ei, double omega);
int mg_resid(const Grid &g, hipStream_t s, double lengthsize, double *res, const double *u, const double *rhs,
             const double *a, const double *b, const double *c);
int mg_rstrct(const Grid &gc, hipStream_t s, double *uc, const double *uf);
int mg_interp(const Grid &gf, hipStream_t s, double *uf, const double *uc);
int mg_addint(const Grid &gf, hipStream_t s, double *uf, const double *uc, double *res);
// split (when given): also the red-black split copy of the coarse a, b, c
// of type split_ct (RB_CT_*), as mg_abcei
int mg_acfromac(const Grid &gc, hipStream_t s, double *a, double *b, double *c, double *ei, const double *a2,
                const double *b2, const double *c2, void *split = nullptr, int split_ct = -1);
// split (when given): also the red-black split copy of a, b, c of type
// split_ct (RB_CT_*), three planes of 2 * half entries
int mg_abcei(const Grid &g, hipStream_t s, double *a, double *b, double *c, double *ei, const double *cc, double r1,
             double r2, void *split = nullptr, int split_ct = RB_CT_F32);
int mg_filterc27(const Grid &g, hipStream_t s, const double *cc, double *cf);
int mg_fill(size_t n, hipStream_t s, double *a, double v);
int op_mglin(Ctx &c);

// tolerance-mode pressure solve of the finest-level system (no reference counterpart)
int op_mg_solve_tol(Ctx &c);

// PLIC functions on batches (device)
int plic_al3d_batch(int n, hipStream_t s, const double *b1, const double *b2, const double *b3, const double *cc,
                    double *out);
int plic_fl3d_batch(int n, hipStream_t s, const double *m1, const double *m2, const double *m3, const double *al,
                    const double *r0, const double *dr0, double *out);

}  // namespace dg
