/*
 * droplet_gpu.h -- C-ABI of libdroplet_gpu.so, the MI355X (gfx950) drop-in
 * for the per-time-step path of farzan-t/3D_Droplet_simulation_2007.
 *
 * Conventions (Fortran-friendly, so the reference's Fortran-77 host can bind
 * them through ISO_C_BINDING interface blocks):
 *  - every scalar argument is passed by reference (const int*, const double*);
 *  - arrays are FP64, column-major (nx,ny,nz) with one ghost layer, exactly
 *    the reference's layout (input.h:33), passed as plain host pointers;
 *  - every function returns 0 on success and -1 on failure; dg_last_error()
 *    then names the failure.  The reference has no error codes (it STOPs);
 *    failures that the reference STOPs on are reported the same way here.
 *  - like the reference's COMMON state, the device state is process-wide and
 *    calls are not re-entrant.  Field state stays resident in HBM between
 *    calls; only dg_get_field/dg_set_field and the host-pointer kernels move
 *    data across PCIe.
 */
#ifndef DROPLET_GPU_H
#define DROPLET_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* field ids for dg_get_field / dg_set_field */
enum {
    DG_F_CC = 0, DG_F_U, DG_F_V, DG_F_W, DG_F_P,
    DG_F_S11, DG_F_S22, DG_F_S33, DG_F_S12, DG_F_S13, DG_F_S23,
    DG_F_RHON, DG_F_MUN, DG_F_MUT, DG_F_KTUR, DG_F_EPS,
    DG_F_A, DG_F_B, DG_F_C, DG_F_EI, DG_F_RHS, DG_F_RES,
    DG_F_PPREV,
    /* CalcH3d's interface heights (/NEWW/), present with hswitch = 1 or
     * surfacet = 5, 6 */
    DG_F_HXMIN, DG_F_HXMAX, DG_F_HYMIN, DG_F_HYMAX, DG_F_HZMIN, DG_F_HZMAX,
    /* /TURBULENCE/ Reynolds stresses of the anisotropic k-epsilon model
     * (uv2p, iturb = 4; present with iturb = 4) and the model coefficient
     * cmu */
    DG_F_U2P, DG_F_V2P, DG_F_W2P, DG_F_UVP, DG_F_UWP, DG_F_VWP,
    DG_F_CMU,
    DG_NFIELDS
};

/* Scalars the host reads or sets between calls: /PHYSICS/ tau, time;
 * /DIFF/ tauh, tauh2 (INTEGER in input.h), tswap; /CYCLING/ NPRE, NTOP,
 * NPOST (input.h:35-60) plus the last step's solver diagnostics. */
typedef struct dg_state {
    double tau, time, tauh;
    int tauh2, tswap, npre, ntop, npost;
    int nx, ny, nz, ng;
    int mg_iters;        /* PCG iterations of the last tolerance-mode solve */
    double mg_relres;    /* its final ||b-Ap||/||b|| */
    double cfl;          /* CFL number computed by the last checktime */
    double epsdiv;       /* h*max|div| of the last adaptive-smoothing check */
} dg_state;

/* ---- lifecycle ------------------------------------------------------- */

/* Replaces initialize (main.f:121-183): readinput (main.f:185-272) of the
 * unchanged inputdata.in format, memoryaloc (main.f:274-322), nondim
 * (main.f:341-446), the droplet IC smooth2 (main.f:5678-5752), bc_c, bc_v,
 * calcrho.  With restart = 1 in inputdata.in, iniget (main.f:698-793)
 * replaces smooth2: u, v, w, cc and the values the reference reads from
 * save.dat, the file in inputdata.in's directory.  Allocates every field in
 * HBM. */
int dg_init(const char *inputdata_path);
/* smooth2 (main.f:5678-5752) generalised to n droplets, replacing the
 * volume fraction and velocity of the resident state (then bc_c, bc_v,
 * calcrho).  params: 7 doubles per droplet in the solver's units (lengths
 * in units of length_SI, on smooth2's grid coordinates (i - 1.5) h;
 * velocities in units of the run's velocity scale): centre x0, y0, z0,
 * radius, liquid velocity u, v, w.  A cell inside one droplet is full, a
 * cell that meets none empty, others count smooth2's 30^3 sub-points in any
 * droplet; velocities are smooth2's mixture rule with the velocity of the
 * cell's droplet (u = u_d cc).  At most 8 droplets may reach into one cell.
 * dg_smooth2_droplet gives the single droplet dg_init placed, in these
 * units: dg_init_droplets with it alone reproduces dg_init's state.  No
 * reference counterpart for n > 1 (the reference has one droplet). */
int dg_init_droplets(const int *n, const double *params);
int dg_smooth2_droplet(double *params);
/* The loop bounds of the run: t0 (0, or save.dat's t0 on a restart) and
 * tmax (t0 + the file's tmax on a restart, main.f:791). */
int dg_run_limits(int *t0, int *tmax);
/* Frees device memory and finishes queued output (the reference has no
 * counterpart; process exit). */
int dg_finalize(void);
/* 0 if the library is usable, -1 after an error. */
int dg_status(void);
const char *dg_last_error(void);

/* ---- the time-step operators (device resident) ------------------------ */
int dg_step(const int *t);          /* loop body main.f:21-71 (printout excluded) */
/* the loop main.f:19-73 for nsteps steps from t0; printout at the start of
 * each step when dg_output_open has opened the output files */
int dg_run(const int *t0, const int *nsteps);
int dg_vofmethod(const int *t);     /* VOFmethod main.f:795-824 */
int dg_kepsmodel(const int *t);     /* kepsmodel main.f:1888-2026 (iturb=3, 4) */
int dg_eddyvisc(void);               /* eddyvisc main.f:1828-1885 (iturb=1, 2) */
int dg_uv2p(void);                   /* uv2p main.f:2032-2155 (iturb=4) */
int dg_checktime(const int *t);     /* checktime main.f:2156-2175 */
int dg_csf(void);                   /* csf main.f:2363-2481 (surfacet=2) */
int dg_calch3d(void);               /* CalcH3d main.f:1312-1810 (hswitch=1) */
int dg_sgip(void);                  /* SGIP main.f:2835-3027 (surfacet=5) */
int dg_sgip2(void);                 /* SGIP2 main.f:3029-3166 (surfacet=6) */
int dg_solvelpres(const int *t);    /* solvelpres main.f:3236-3439 */
int dg_calcrho(void);               /* calcrho main.f:2177-2190 */
/* CFL() main.f:4848-4881: max(|u|,|v|,|w|) over the interior times tauh,
 * on the resident velocity; NaN velocities are an error */
int dg_cfl(double *cfl);
int dg_mglin(void);                 /* mglin main.f:3645-3763 on the resident finest level */

/* ---- state exchange (replaces reading/writing the COMMON blocks) ------ */
int dg_get_field(const int *id, double *host);
int dg_set_field(const int *id, const double *host);
int dg_get_state(dg_state *out);
int dg_set_state(const dg_state *in);   /* sets tau,time,tauh,tauh2,tswap,npre,ntop,npost */
/* Clips the resident volume fraction to [0,1] (a deviation from the
 * reference: smooth2 can leave 1+4e-13, which sends the reference's AL3D,
 * main.f:4562-4634, to NaN). */
int dg_clip_cc(void);

/* Pressure-solver selection (no reference counterpart):
 * mode 0 = the reference's mglin schedule (default, bit-parity path),
 * mode 1 = FP64 PCG with an FP64 V-cycle preconditioner to ||r||/||b|| < tol,
 * mode 2 = FP64 PCG with an FP32 V-cycle preconditioner to ||r||/||b|| < tol.
 * guess: 0 zero, 1 previous step's pressure (dropped for zero when its
 * residual is no smaller than ||b||), 2 learned (dg_guess_*) from zero. */
int dg_set_solver(const int *mode, const double *tol, const int *nu1, const int *nu2, const int *maxit,
                  const int *guess);
/* Tolerance-mode smoother (no reference counterpart): SOR weight omega of
 * the symmetric red-black passes (1, the default: Gauss-Seidel) and the
 * sweep count nu_coarse before and after the coarse correction on every
 * level below the finest (0, the default: nu1 / nu2).  Applies to both
 * precisions, the resident state and later dg_init / dg_poisson_solve. */
int dg_set_smoother(const double *omega, const int *nu_coarse);
/* 1 (default): outflow sums in the reference's sequential order
 * (main.f:4123-4131); 0: parallel tree sums (different rounding). */
int dg_set_ordered_sums(const int *on);
/* 1: a partly filled cell whose Youngs normal vanishes (no filled
 * neighbour, or a symmetric neighbourhood) is advected with a plane normal
 * to the sweep direction; the reference's AL3D/FL3D evaluate 0/0 there and
 * the run turns to NaN (swp main.f:848-1000).  0 (default): the reference's
 * arithmetic. */
int dg_set_vof_guard(const int *on);
/* Tolerance-mode execution switches.  graphs = 1 (default): each PCG
 * iteration replayed as a captured HIP graph, 0 launched kernel by kernel
 * (bit-identical).  split = 1 (default):
 * mode 2's FP32 V-cycle keeps each level above the coarsest in red-black
 * split storage with FP16 face coefficients when the densities allow
 * (different rounding of the preconditioner, same tolerance reached);
 * 2: split storage, FP32 coefficients (bit-identical to 0); 3: split, FP16
 * coefficients always; 0: natural layout, FP32 coefficients.  tail (split
 * storage only; bit-identical either way): the levels of at most 4096 cells
 * and the coarsest solve run as 0 one launch per operation, 1 one
 * single-workgroup kernel in global memory, 2 (default) the same kernel on
 * LDS copies of those levels when they fit.  Applies to the resident state
 * and to later dg_init / dg_poisson_solve calls. */
int dg_set_tuning(const int *graphs, const int *split, const int *tail);

/* Tile smoother of the tolerance-mode V-cycle (no reference counterpart;
 * split storage only): on = 2 runs each level strictly between the finest
 * level and the tail (levels of more than 4096 cells), on = 1 those of them
 * with at most 2^18 cells, as tiles of 16x8x8
 * cells, all of a visit's pre- or post-smoothing sweeps in one launch: each
 * tile smooths itself grown by one cell in LDS with the cells around that
 * held at their values from before the sweeps, and writes back its own cells
 * (restricted additive Schwarz, red-black SOR inside; the prolongation is
 * folded into the post-smoothing launch).  nu: the sweeps before and after
 * the coarse correction on those levels (0: the levels' own count, the
 * smoother's nu_coarse).  A different preconditioner from the global
 * passes (on = 0, the default), the same tolerance reached.  Applies
 * to the resident state and to later dg_init / dg_poisson_solve calls. */
int dg_set_tile_smoother(const int *on, const int *nu);

/* Look-ahead of the tolerance-mode PCG loop: the host queues iteration n+1
 * before iteration n's residual reaches it while the last residual it has
 * seen exceeds ratio * tol (default 1e3; <= 0 never), so the GPU does not
 * idle between the two.  Bit-identical either way: an iteration queued
 * after convergence is skipped by the device's own stopping test (its
 * V-cycle still runs) and changes nothing the solve returns.  Applies to
 * the resident state and to later dg_init / dg_poisson_solve calls. */
int dg_set_lookahead(const double *ratio);

/* ---- output, checkpoint and restart ------------------------------------ */
/* openfiles (main.f:105-118) in directory dir ("" or NULL: the current
 * one): creates outputdata.out, spurious.plt, normv.plt, hysteresis.plt and
 * velocity.plt with their header lines, and starts the writer thread.
 * echo_stdout = 1 also writes printout's screen lines to stdout. */
int dg_output_open(const char *dir, const int *echo_stdout);
/* printout(t, cc, p) (main.f:5020-5100) at the start of step t, writing the
 * reference's files in its formats: every tprint steps velocityc's record
 * (velocity.plt), the diagnostics lines (outputdata.out, and stdout when
 * echoed) and the spurious.plt and normv.plt records; every ntecplot steps
 * mainNNNN.plt (printtecp, main.f:5106-5153); every 100 steps
 * pressureNNNN.plt; at t = tsave the restart records (savedata) into
 * save.dat -- the file iniget reads (the reference writes them to unit 2
 * after opening saved.dat on unit 3, main.f:5418-5419).  The fields are
 * snapshotted on the device in stream order and downloaded and written by
 * the writer thread: the call returns before the files are written. */
int dg_printout(const int *t);
/* Waits until every queued printout is written and flushes the files;
 * returns -1 with the writer's first error. */
int dg_output_flush(void);
/* Finishes queued output, closes the files and stops the writer thread. */
int dg_output_close(void);
/* The writer's counters since the process started: printouts queued,
 * bytes downloaded, milliseconds dg_printout waited for a free snapshot
 * slot, milliseconds the writer thread spent on files. */
int dg_output_stats(int *jobs, double *bytes, double *wait_ms, double *write_ms);
/* savedata (main.f:5414-5467): the restart records of the current u, v, w,
 * cc and the run's values into the file `path` (Fortran unformatted
 * sequential records, the layout iniget reads). */
int dg_savedata(const char *path);
/* The full resident state -- every scalar, every field that carries over
 * between steps and every level's multigrid heap arrays -- into `path`,
 * and back into a run of the same grid (no reference counterpart: the
 * reference's restart file holds u, v, w, cc only).  A run resumed from
 * dg_save_state after step t continues bit-identically with step t + 1. */
int dg_save_state(const char *path);
int dg_load_state(const char *path);
/* Host formatting, no GPU (no reference counterpart): x as flang writes it
 * list-directed (mode 0) or with the reference's E15.7 (mode 1: the
 * writers' form; mode 2: the same through printf's correctly rounded
 * digits, a check of mode 1), into buf of len bytes, NUL-terminated. */
int dg_format_real(const double *x, const int *mode, char *buf, const int *len);

/* ---- host-pointer kernels with the reference's argument lists --------- */
/* Each uploads its arrays, runs on the GPU and downloads the outputs. */
int dg_relax(double *u, const double *lengthsize, const double *rhs, const double *a, const double *b,
             const double *c, const double *ei, const int *ibc, const int *nx, const int *ny, const int *nz,
             const double *omega);                                           /* relax main.f:3956 */
int dg_resid(const double *lengthsize, double *res, const double *u, const double *rhs, const double *a,
             const double *b, const double *c, const int *ibc, const int *nx, const int *ny,
             const int *nz);                                                 /* resid main.f:3902 */
int dg_rstrct(double *uc, const double *uf, const int *ibc, const int *ncx, const int *ncy,
              const int *ncz);                                               /* rstrct main.f:3765 */
int dg_interp(double *uf, const double *uc, const int *ibc, const int *nfx, const int *nfy,
              const int *nfz);                                               /* interp main.f:3866 */
int dg_addint(double *uf, const double *uc, double *res, const int *ibc, const int *nfx, const int *nfy,
              const int *nfz);                                               /* addint main.f:3934 */
int dg_acfromac(double *a, double *b, double *c, double *ei, const double *a2, const double *b2,
                const double *c2, const int *nx, const int *ny, const int *nz); /* acfromac main.f:3793 */
int dg_slvsml(double *u, const double *rhs, const double *a, const double *b, const double *c, const double *ei,
              const int *ibc, const int *nx, const int *ny, const int *nz, const int *ntop, const double *omega,
              const double *lengthsize);                                     /* slvsml main.f:3846 */
int dg_abcei(double *a, double *b, double *c, double *ei, const double *cc, const double *r1, const double *r2,
             const int *ibc, const int *nx, const int *ny, const int *nz);   /* abcei main.f:3582 */
int dg_bc_scalar(double *scal, const int *ibc, const int *nx, const int *ny, const int *nz); /* main.f:4298 */
int dg_filterc27(const double *cc, double *cfilt, const int *nx, const int *ny, const int *nz); /* main.f:5473 */
/* AL3D (main.f:4562) and FL3D (main.f:4639) over n independent inputs */
int dg_al3d_batch(const int *n, const double *b1, const double *b2, const double *b3, const double *cc,
                  double *out);
int dg_fl3d_batch(const int *n, const double *m1, const double *m2, const double *m3, const double *alpha,
                  const double *r0, const double *dr0, double *out);

/* Tolerance-mode solve of resid's operator (main.f:3902-3931) on host arrays
 * (no reference counterpart): p is the initial guess on entry and the
 * solution on return (ghosts filled as bc_scalar).  a,b,c as built by abcei. */
int dg_poisson_solve(const int *nx, const int *ny, const int *nz, const double *lengthsize, const double *rhs,
                     const double *a, const double *b, const double *c, double *p, const int *mode,
                     const double *tol, const int *maxit, int *iters, double *relres);

/* ---- learned initial guess (no reference counterpart) ------------------ */
/* A 3D encoder-decoder (channels 16/32/64/64, 3x3x3 convolutions, BF16
 * weights and activations, FP32 accumulation on the matrix cores) on the
 * next-coarser multigrid level predicts the direction of the step's
 * pressure from (log10 mean face coefficient / 3, asinh(b/rms(b)) and
 * asinh(100 b/rms(b))/10 of the right-hand side b, previous pressure /
 * rms); captured targets (the converged pressure) are scaled to unit rms
 * per sample.  With dg_set_solver guess = 2 the tolerance-mode solve starts
 * from alpha * prolong(prediction), alpha = (d.b)/(d.Ad), and falls back to
 * the zero guess when that does not lower |b - A p|.  Network extents
 * ((n-2)/2 per axis) must be multiples of 8.
 *
 * Parameters: one FP32 vector, per layer [coutp][27][cinp] weights then
 * [coutp] biases (padded rows and columns zero). */
int dg_guess_nparams(int *n);
int dg_guess_init_random(const int *seed);
int dg_guess_set_params(const float *params, const int *n);
int dg_guess_get_params(float *params, const int *n);
/* network on host arrays: feat [nz][ny][nx][4] -> out [nz][ny][nx] */
int dg_guess_forward(const int *nx, const int *ny, const int *nz, const float *feat, float *out);
/* loss 0.5*mean((out-target)^2) and its gradient for every parameter */
int dg_guess_gradient(const int *nx, const int *ny, const int *nz, const float *feat, const float *target,
                      double *loss, float *grad);
/* one Adam step (b1 0.9, b2 0.999, eps 1e-8) on one sample */
int dg_guess_train_sample(const int *nx, const int *ny, const int *nz, const float *feat, const float *target,
                          const float *lr, double *loss);
/* Sample store in device memory: with capacity > 0 every tolerance-mode
 * solve stores its network input and target (ring buffer); 0 frees it. */
int dg_guess_capture(const int *capacity);
int dg_guess_store_count(int *n);
/* sample i: feat [V][4] and target [V] (either may be NULL), dims[3] */
int dg_guess_store_get(const int *i, float *feat, float *target, int *dims);
/* epochs of Adam over the stored samples, reshuffled from seed each epoch */
int dg_guess_train(const int *epochs, const float *lr, const int *seed, double *loss_per_epoch);
/* solves that took / rejected the learned step, solves the controller ran
 * without the network, the last step length and the mean of
 * |r|^2 before / after the learned step over the solves that ran it */
int dg_guess_stats(int *accepted, int *rejected, int *skipped, double *last_alpha, double *mean_gain);
/* Fallback controller: a learned step that divides |r|^2 by less than
 * min_gain sends the next probe_every * 2^(w-1) solves to the zero guess
 * alone, w the number of such steps in a row (capped at 16 probe_every; a
 * step that reaches min_gain resets w).  probe_every 0: the network runs
 * on every solve.  Default 8, 4. */
int dg_guess_set_controller(const int *probe_every, const double *min_gain);
/* on = 1: each solve runs from the zero, previous and learned guesses in
 * turn (same system); dg_guess_eval_iters returns the three counts of the
 * last solve, whose result is the learned solve's. */
int dg_guess_eval(const int *on);
int dg_guess_eval_iters(int *iters);
int dg_guess_release(void);
/* single layers on host arrays (operands rounded to BF16; x [V][cinp],
 * w [coutp][27][cinp], y [V][coutp] (values of the BF16 result)) */
int dg_conv3d_fwd(const int *nx, const int *ny, const int *nz, const float *x, const int *cinp, const float *w,
                  const float *bias, const int *cout, const int *coutp, const int *relu, float *y);
/* dx [V][cinp] = adjoint of the convolution applied to dy [V][coutp] */
int dg_conv3d_dgrad(const int *nx, const int *ny, const int *nz, const float *dy, const int *coutp, const float *w,
                    const int *cinp, float *dx);
/* gw [coutp][27][cinp] and gb [coutp] (FP32) from x [V][cinp], dy [V][coutp] */
int dg_conv3d_wgrad(const int *nx, const int *ny, const int *nz, const float *x, const int *cinp, const float *dy,
                    const int *coutp, float *gw, float *gb);

/* ---- the reference's own call forms (Fortran-77, implicit interfaces) --- */
/* Definitions of the time step's procedures under their Fortran names, with
 * their argument lists, for a program built from main.f: take these units
 * out of main.f and link it against this library (oracle/link_dropin.py
 * does exactly that).  State lives in the program's COMMON blocks
 * (input.h:19-95), which the library binds through their symbols
 * (memory_, pointers_, surfcom_, turbulence_, dimension_, diff_, cycling_,
 * times_, dinem_, physics_).  The first call uploads every field and
 * multigrid heap slot; afterwards the device copies are authoritative, an
 * array argument names its resident copy by address (z(icc(ng)), u, ...;
 * any other address is an error), the scalars tau, time, tauh, tauh2,
 * tswap, NPRE, NTOP, NPOST are read from COMMON before each call and written
 * back after it, and fields return to host memory through
 * dg_sync_to_host_.  On an error these procedures print the message and
 * stop the program with exit status 2, as the reference STOPs, unless
 * dg_set_stop_on_error_(0) was called (then dg_status_ returns -1). */
void vofmethod_(const int *t);                                   /* VOFmethod(t)        main.f:795 */
void kepsmodel_(const int *t, double *kturnew, double *epsnew);  /* kepsmodel(t,..)     main.f:1888 */
void checktime_(const int *t);                                   /* checktime(t)        main.f:2156 */
double cfl_(void);                                               /* CFL()               main.f:4848 */
void csf_(double *ccc, double *w11, double *w22, double *w33);   /* csf(ccc,w11,w22,w33) main.f:2363 */
void solvelpres_(double *cc, double *w1, double *w2, double *w3, double *a, double *b, double *c, double *ei,
                 const int *t);                                  /* solvelpres(...,t)   main.f:3236 */
void calcrho_(double *cc);                                       /* calcrho(cc)         main.f:2177 */
void calch3d_(double *cc);                                       /* CalcH3d(cc)         main.f:1312 */
void eddyvisc_(double *cc, double *ufilt);                       /* eddyvisc(cc,ufilt)  main.f:1828 */
void uv2p_(void);                                                /* uv2p()              main.f:2032 */
void sgip_(double *ccc, double *w11, double *w22, double *w33);  /* SGIP(ccc,...)       main.f:2835 */
void sgip2_(double *ccc, double *w11, double *w22, double *w33); /* SGIP2(ccc,...)      main.f:3029 */
void bc_c_(double *cc);                                          /* bc_c(cc)            main.f:4331 */
void mglin_(const double *r1, const double *r2);                 /* mglin(r1,r2)        main.f:3645 */
void swp_(double *us, double *cc, double *vof1, double *vof2, double *vof3, const int *d); /* swp main.f:848 */
void convective_(double *w1, double *w2, double *w3);            /* convective(w1,w2,w3) main.f:3441 */
/* The fine-grained routines under their own names, on host arrays (upload,
 * run, download; no COMMON needed), with the argument lists of the dg_*
 * host-pointer kernels above.  For unit drop-in and tests. */
void relax_(double *u, const double *lengthsize, const double *rhs, const double *a, const double *b,
            const double *c, const double *ei, const int *ibc, const int *nx, const int *ny, const int *nz,
            const double *omega);                                        /* relax      main.f:3956 */
void resid_(const double *lengthsize, double *res, const double *u, const double *rhs, const double *a,
            const double *b, const double *c, const int *ibc, const int *nx, const int *ny,
            const int *nz);                                              /* resid      main.f:3902 */
void rstrct_(double *uc, const double *uf, const int *ibc, const int *ncx, const int *ncy,
             const int *ncz);                                            /* rstrct     main.f:3765 */
void interp_(double *uf, const double *uc, const int *ibc, const int *nfx, const int *nfy,
             const int *nfz);                                            /* interp     main.f:3866 */
void addint_(double *uf, const double *uc, double *res, const int *ibc, const int *nfx, const int *nfy,
             const int *nfz);                                            /* addint     main.f:3934 */
void acfromac_(double *a, double *b, double *c, double *ei, const double *a2, const double *b2,
               const double *c2, const int *nx, const int *ny, const int *nz); /* acfromac main.f:3793 */
void slvsml_(double *u, const double *rhs, const double *a, const double *b, const double *c, const double *ei,
             const int *ibc, const int *nx, const int *ny, const int *nz, const int *ntop, const double *omega,
             const double *lengthsize);                                  /* slvsml     main.f:3846 */
void abcei_(double *a, double *b, double *c, double *ei, const double *cc, const double *r1, const double *r2,
            const int *ibc, const int *nx, const int *ny, const int *nz); /* abcei     main.f:3582 */
void bc_scalar_(double *scal, const int *ibc, const int *nx, const int *ny, const int *nz); /* main.f:4298 */
void filterc27_(const double *cc, double *cfilt, const int *nx, const int *ny, const int *nz); /* main.f:5473 */
double al3d_(const double *b1, const double *b2, const double *b3, const double *cc); /* AL3D main.f:4562 */
double fl3d_(const double *m1, const double *m2, const double *m3, const double *alpha, const double *r0,
             const double *dr0);                                         /* FL3D       main.f:4639 */
/* Lifecycle of the call forms (no reference counterpart).  dg_attach_:
 * after the program's initialize (main.f:15), bind COMMON and upload
 * (otherwise done by the first call).  dg_init_input_: for a program
 * without the reference's initialize, the library reads ./inputdata.in and
 * sets up the run (as dg_init) and writes the result into COMMON (scalars,
 * /pointers/ with MYMALLOC's bookkeeping, every field).  dg_sync_to_host_ /
 * dg_sync_to_device_(mask): bit id for each DG_F_* field id below
 * DG_F_U2P that has a host array, bit 29 for cmu, bit 31 for the six
 * Reynolds-stress arrays, bit 30 for every level's heap slots, -1 for all. */
void dg_attach_(void);
void dg_init_input_(void);
void dg_sync_to_host_(const int *mask);
void dg_sync_to_device_(const int *mask);
int dg_status_(void);
void dg_set_stop_on_error_(const int *on);
void dg_finalize_(void);

/* ---- no-allocation entry point ------------------------------------------ */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): CFL() of main.f:4848-4881, as dg_cfl computes it, on
 * velocity arrays the caller holds in device memory, nx*ny*nz doubles each
 * in the reference layout.  The result, max(|u|,|v|,|w|) over the interior
 * times tauh, is written to the device double *cfl; the call enqueues its
 * kernels on `stream` (a hipStream_t, NULL for the null stream) and returns
 * without waiting.  Allocates nothing: the reduction's partials live in a
 * fixed 8200-byte device array that every GPU the library runs on holds
 * until the process ends, shared by all calls on that GPU (calls that
 * overlap on one GPU race).  Grids above DG_CFL_NOALLOC_MAX_CELLS cells
 * (the kernels' 32-bit cell indices) are rejected with an error. */
#define DG_CFL_NOALLOC_MAX_CELLS 2147483647LL
int dg_cfl_noalloc(const double *u, const double *v, const double *w, const int *nx, const int *ny, const int *nz,
                   const double *tauh, double *cfl, void *stream);

/* ---- measurement (no reference counterpart) ---------------------------- */
/* Times every launch of one kernel family with HIP events recorded on the
 * library's stream until dg_profile_stop, which returns the summed kernel
 * time, the launch count and the summed algorithmic bytes of those launches.
 * kernel ids: 1 tolerance-mode Gauss-Seidel colour pass, 2 its residual,
 * 3 PCG direction + A.p, 4 PCG update, 5 reference relax colour pass,
 * 6 fused two-colour smoother sweep, 7 finest-level colour pass of the
 * red-black split FP32 V-cycle, 8 BF16 MFMA convolution forward of the guess
 * network (total_bytes then holds the launches' algorithmic FLOPs). */
int dg_profile_start(const int *kernel_id);
int dg_profile_stop(double *total_ms, int *launches, double *total_bytes);

#ifdef __cplusplus
}
#endif
#endif
