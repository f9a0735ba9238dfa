// dg_api.hip -- C-ABI of libdroplet_gpu.so: lifecycle, host set-up
// (readinput / memoryaloc / nondim of the reference), the time-step
// operators and the host-pointer kernel entry points.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "droplet_gpu.h"
#include "dg_ctx.h"
#include "dg_prof.h"
#include "dg_util.h"
#include "dg_resident.h"

namespace dg {

static char g_err[1024] = "";
static int g_status = 0;
static thread_local std::string *t_err_sink = nullptr;

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    if (t_err_sink) {
        // a worker thread's error: kept for the thread (the first one)
        char buf[1024];
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        if (t_err_sink->empty()) *t_err_sink = buf;
        return;
    }
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    g_status = -1;
}
void set_error_sink(std::string *sink) { t_err_sink = sink; }
const char *last_error() { return g_err; }
int status() { return g_status; }

Ctx *ctx()
{
    static Ctx c;
    return &c;
}

Prof &prof()
{
    static Prof p;
    return p;
}

static bool g_ready = false;
static int g_graphs_default = 1, g_split_default = 1, g_tail_default = 2;
static double g_omega_default = 1.0;  // tolerance-mode smoother (dg_set_smoother)
static int g_nuc_default = 0;
static int g_tiles_default = 0, g_tile_nu_default = 0;  // tile smoother (dg_set_tile_smoother)
static double g_lookahead_default = 1e3;  // dg_set_lookahead

// ---- allocation -----------------------------------------------------------

static int dalloc(double **p, size_t n)
{
    DG_HIP(hipMalloc(p, n * sizeof(double)));
    DG_HIP(hipMemset(*p, 0, n * sizeof(double)));
    return 0;
}

// level sizes: memoryaloc (main.f:274-322) halves n -> n/2+1 while the
// smallest extent exceeds 4.  Extents of the form m*2^k+2 are accepted for
// the tolerance-mode solver (halving stops when an interior extent turns odd).
static int plan_levels(int nx, int ny, int nz, bool strict, int dims[MAXLEV][3], int *ng_out)
{
    int nn = nx < ny ? nx : ny;
    if (nz < nn) nn = nz;
    if (strict) {
        int j = nn - 2, ng = 1;
        while (j > 2) { j = j / 2; ng++; }
        if ((1 << ng) + 2 != nn) {
            set_error("STOP: nn .ne. 2**ng + 2 (nn=%d)", nn);
            return -1;
        }
    }
    int lx = nx, ly = ny, lz = nz, cnt = 0;
    int tmp[MAXLEV][3];
    for (;;) {
        if (cnt >= MAXLEV) { set_error("too many multigrid levels"); return -1; }
        tmp[cnt][0] = lx; tmp[cnt][1] = ly; tmp[cnt][2] = lz;
        cnt++;
        int m = lx < ly ? lx : ly;
        if (lz < m) m = lz;
        if (m <= 4) break;
        if (((lx - 2) & 1) || ((ly - 2) & 1) || ((lz - 2) & 1)) break;
        lx = lx / 2 + 1; ly = ly / 2 + 1; lz = lz / 2 + 1;
    }
    for (int l = 0; l < cnt; l++)
        for (int d = 0; d < 3; d++) dims[l][d] = tmp[cnt - 1 - l][d];  // dims[0] = coarsest
    *ng_out = cnt;
    return 0;
}

static int alloc_levels(Ctx &c, int nx, int ny, int nz, bool strict)
{
    int dims[MAXLEV][3], ng;
    if (plan_levels(nx, ny, nz, strict, dims, &ng)) return -1;
    c.ng = ng;
    {
        // the reference's form nn = 2**ng + 2, which its FMG schedule needs
        int nn = nx < ny ? nx : ny;
        if (nz < nn) nn = nz;
        int j = nn - 2, ngr = 1;
        while (j > 2) { j = j / 2; ngr++; }
        c.ref_grid = (1 << ngr) + 2 == nn && ngr == ng;
    }
    for (int l = 0; l < ng; l++) {
        Level &L = c.lev[l];
        L.g = make_grid(dims[l][0], dims[l][1], dims[l][2]);
        const size_t n = L.g.n();
        if (dalloc(&L.rhs, n) || dalloc(&L.res, n) || dalloc(&L.u, n) || dalloc(&L.cc, n) || dalloc(&L.a, n) ||
            dalloc(&L.b, n) || dalloc(&L.c, n) || dalloc(&L.ei, n))
            return -1;
    }
    c.g = c.lev[ng - 1].g;
    // reduction partials: RED_BLOCKS pairs or two per interior block
    const dim3 ig = interior_grid(c.g);
    // two values per interior block, then room for the second-stage sums
    size_t np = 2 * (size_t)ig.x * ig.y * ig.z;
    if (np < 2 * (size_t)RED_BLOCKS) np = 2 * RED_BLOCKS;
    np += np / 512 + 64;
    if (dalloc(&c.red_partial, np)) return -1;
    if (dalloc(&c.dscal, 64)) return -1;
    DG_HIP(hipHostMalloc((void **)&c.hscal, 64 * sizeof(double), hipHostMallocDefault));
    memset(c.hscal, 0, 64 * sizeof(double));
    DG_HIP(hipHostMalloc((void **)&c.hpoll, 8 * sizeof(double), hipHostMallocMapped | hipHostMallocCoherent));
    memset(c.hpoll, 0, 8 * sizeof(double));
    DG_HIP(hipHostGetDevicePointer((void **)&c.dpoll, c.hpoll, 0));
    if (!c.stream) DG_HIP(hipStreamCreate(&c.stream));
    return 0;
}

static int alloc_fields(Ctx &c)
{
    const size_t n = c.g.n();
    double **arrs[] = {&c.u, &c.v, &c.w, &c.s11, &c.s22, &c.s33, &c.s12, &c.s13, &c.s23, &c.rhon,
                       &c.mun, &c.mut, &c.ktur, &c.eps, &c.cmu, &c.ccf, &c.pprev};
    for (double **p : arrs)
        if (dalloc(p, n)) return -1;
    return dalloc(&c.lmix, 2 * (size_t)c.g.nx);  // lmix, then eddyvisc's per-station scratch
}

// CalcH3d's six height arrays, zeroed (the reference's COMMON starts at 0
// and its ghost cells are never written)
int alloc_reynolds(Ctx &c)
{
    if (c.u2p) return 0;
    const size_t n = c.g.n();
    double **arrs[] = {&c.u2p, &c.v2p, &c.w2p, &c.uvp, &c.uwp, &c.vwp};
    for (double **p : arrs)
        if (dalloc(p, n)) return -1;
    return 0;
}

int alloc_heights(Ctx &c)
{
    if (c.hxmin) return 0;
    const size_t n = c.g.n();
    double **arrs[] = {&c.hxmin, &c.hxmax, &c.hymin, &c.hymax, &c.hzmin, &c.hzmax};
    for (double **p : arrs)
        if (dalloc(p, n)) return -1;
    return 0;
}

static void free_ctx(Ctx &c)
{
    for (int l = 0; l < c.ng; l++) {
        Level &L = c.lev[l];
        void *ps[] = {L.rhs, L.res, L.u, L.cc, L.a, L.b, L.c, L.ei, L.a32, L.b32, L.c32,
                      L.u32, L.f32, L.r32, L.u64, L.f64, L.r64, L.rb_u, L.rb_f, L.rb_w, L.rb_u64,
                      L.rb_f64, L.rb_w64, L.rb_coef};
        for (void *p : ps)
            if (p) (void)hipFree(p);
        L = Level();
    }
    void *fs[] = {c.u, c.v, c.w, c.s11, c.s22, c.s33, c.s12, c.s13, c.s23, c.rhon, c.mun, c.mut, c.ktur,
                  c.eps, c.cmu, c.ccf, c.pprev, c.pcg_r, c.pcg_p, c.pcg_q, c.red_partial, c.dscal, c.coarse_M,
                  c.hxmin, c.hxmax, c.hymin, c.hymax, c.hzmin, c.hzmax, c.u2p, c.v2p, c.w2p, c.uvp, c.uwp,
                  c.vwp, c.lmix};
    for (void *p : fs)
        if (p) (void)hipFree(p);
    for (int q = 0; q < 2; q++)
        if (c.pcg_graph[q]) (void)hipGraphExecDestroy(c.pcg_graph[q]);
    if (c.hscal) (void)hipHostFree(c.hscal);
    if (c.hpoll) (void)hipHostFree(c.hpoll);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    c = Ctx();
}

// ---- readinput (main.f:185-272): first token of each of the 38 lines ----

int read_input(const char *path, Input &in)
{
    FILE *f = fopen(path, "r");
    if (!f) {
        set_error("cannot open %s", path);
        return -1;
    }
    double v[38];
    char line[1024];
    for (int q = 0; q < 38; q++) {
        if (!fgets(line, sizeof line, f)) {
            fclose(f);
            set_error("%s: missing value on line %d", path, q + 1);
            return -1;
        }
        char *p = line;
        while (*p == ' ' || *p == '\t') p++;
        char *end = nullptr;
        v[q] = strtod(p, &end);
        if (end == p) {
            fclose(f);
            set_error("%s: no number on line %d", path, q + 1);
            return -1;
        }
    }
    fclose(f);
    int q = 0;
    in.nx = (int)v[q++]; in.ny = (int)v[q++]; in.nz = (int)v[q++]; in.tmax = (int)v[q++];
    in.nfilter = (int)v[q++]; in.cycles = (int)v[q++]; in.omega = v[q++];
    in.npre = (int)v[q++]; in.ntop = (int)v[q++]; in.npost = (int)v[q++]; in.restart = (int)v[q++];
    in.maxcfl = v[q++]; in.iturb = (int)v[q++]; in.convectype = (int)v[q++]; in.hswitch = (int)v[q++];
    in.xvel1 = v[q++]; in.yvel1 = v[q++]; in.zvel1 = v[q++];
    in.xvel2 = v[q++]; in.yvel2 = v[q++]; in.zvel2 = v[q++];
    in.rho2 = v[q++]; in.rho1 = v[q++]; in.sigma = v[q++]; in.visc2 = v[q++]; in.visc1 = v[q++];
    in.grav = v[q++]; in.grav1 = v[q++]; in.icase = (int)v[q++]; in.ibc = (int)v[q++];
    in.length = v[q++]; in.radius = v[q++]; in.y0 = v[q++]; in.z0 = v[q++];
    in.tprint = (int)v[q++]; in.ntecplot = (int)v[q++]; in.tsave = (int)v[q++]; in.surfacet = (int)v[q++];
    return 0;
}

// nondim (main.f:341-446): host arithmetic in the reference's order
void nondim(const Input &in, Scalars &s)
{
    const double L = in.length;
    s.lengthscale = L;
    double vs = sqrt((in.xvel1 - in.xvel2) * (in.xvel1 - in.xvel2) + (in.yvel1 - in.yvel2) * (in.yvel1 - in.yvel2) +
                     (in.zvel1 - in.zvel2) * (in.zvel1 - in.zvel2));
    double cdrag = 1.0;
    const double vel2 = sqrt(in.xvel2 * in.xvel2 + in.yvel2 * in.yvel2 + in.zvel2 * in.zvel2);
    const double gtot = sqrt(in.grav * in.grav + in.grav1 * in.grav1);
    if (vs == 0.0 && gtot != 0.0) {
        vs = sqrt((fabs(gtot) * L) * fabs(in.rho1 - in.rho2) / in.rho2 * (1.0 / cdrag));
        const double re2 = vel2 * L * in.rho2 / in.visc2;
        if (re2 <= 10.0 && vs == 0.0) cdrag = 10.0 / re2;
        vs = sqrt((fabs(gtot) * L) * fabs(in.rho1 - in.rho2) / in.rho2 * (1.0 / cdrag));
    }
    if (vs == 0.0) vs = 1.0;
    s.vscale = vs;
    s.timescale = L / vs;
    s.lengthsize = in.length / L;
    s.h = s.lengthsize / (double)(in.nx - 2);
    s.time = 0.0;
    s.tau_si = in.maxcfl * DG_F4(0.8) * (in.length / (double)(in.nx - 2)) / vs;
    s.tau = s.tau_si / s.timescale;
    s.tauh = s.tau / s.h;
    s.tauh2 = (int)(s.tau / (s.h * s.h));
    s.tswap = 0;
    s.radius = in.radius / L;
    s.xvel1 = in.xvel1 / vs; s.yvel1 = in.yvel1 / vs; s.zvel1 = in.zvel1 / vs;
    s.xvel2 = in.xvel2 / vs; s.yvel2 = in.yvel2 / vs; s.zvel2 = in.zvel2 / vs;
    s.y0 = (double)(in.ny - 2) * s.h / 2.0;
    s.z0 = (double)(in.nz - 2) * s.h / 2.0;
    s.rho1 = 1.0;
    s.rho2 = in.rho2 / in.rho1;
    double denom = in.rho1 * (L * L) / s.timescale;
    s.mu1 = in.visc1 / denom;
    s.mu2 = in.visc2 / denom;
    denom = in.rho1 * (L * L * L) / (s.timescale * s.timescale);
    s.sigma = in.sigma / denom;
    denom = L / (s.timescale * s.timescale);
    s.g = in.grav / denom;
    s.g1 = in.grav1 / denom;
    // smooth2's droplet centre (main.f:5687-5690)
    if (in.z0 != 1000.0) s.z0 = in.z0 / in.length;
    if (in.y0 != 1000.0) s.y0 = in.y0 / in.length;
    s.y0 = s.y0 + s.h / 2.0;
    s.maxcfl = in.maxcfl;
    s.omega = in.omega;
    s.cycles = in.cycles; s.npre = in.npre; s.ntop = in.ntop; s.npost = in.npost;
    s.iturb = in.iturb; s.convectype = in.convectype; s.surfacet = in.surfacet; s.hswitch = in.hswitch;
    s.tprint = in.tprint;
}

static double *field_ptr(Ctx &c, int id)
{
    Level &L = c.lev[c.ng - 1];
    switch (id) {
    case DG_F_CC: return L.cc;
    case DG_F_U: return c.u;
    case DG_F_V: return c.v;
    case DG_F_W: return c.w;
    case DG_F_P: return L.u;
    case DG_F_S11: return c.s11;
    case DG_F_S22: return c.s22;
    case DG_F_S33: return c.s33;
    case DG_F_S12: return c.s12;
    case DG_F_S13: return c.s13;
    case DG_F_S23: return c.s23;
    case DG_F_RHON: return c.rhon;
    case DG_F_MUN: return c.mun;
    case DG_F_MUT: return c.mut;
    case DG_F_KTUR: return c.ktur;
    case DG_F_EPS: return c.eps;
    case DG_F_A: return L.a;
    case DG_F_B: return L.b;
    case DG_F_C: return L.c;
    case DG_F_EI: return L.ei;
    case DG_F_RHS: return L.rhs;
    case DG_F_RES: return L.res;
    case DG_F_PPREV: return c.pprev;
    case DG_F_HXMIN: return c.hxmin;
    case DG_F_HXMAX: return c.hxmax;
    case DG_F_HYMIN: return c.hymin;
    case DG_F_HYMAX: return c.hymax;
    case DG_F_HZMIN: return c.hzmin;
    case DG_F_HZMAX: return c.hzmax;
    case DG_F_U2P: return c.u2p;
    case DG_F_V2P: return c.v2p;
    case DG_F_W2P: return c.w2p;
    case DG_F_UVP: return c.uvp;
    case DG_F_UWP: return c.uwp;
    case DG_F_VWP: return c.vwp;
    case DG_F_CMU: return c.cmu;
    default: return nullptr;
    }
}

static int require_ready()
{
    if (!g_ready) {
        set_error("dg_init has not been called");
        return -1;
    }
    return 0;
}

// ---- scoped device buffers for the host-pointer entry points --------------
struct DevBufs {
    std::vector<double *> ptrs;
    ~DevBufs()
    {
        for (double *p : ptrs) (void)hipFree(p);
    }
    double *up(const double *h, size_t n)
    {
        double *d = nullptr;
        if (hipMalloc(&d, n * sizeof(double)) != hipSuccess) return nullptr;
        ptrs.push_back(d);
        if (h && hipMemcpy(d, h, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
        if (!h && hipMemset(d, 0, n * sizeof(double)) != hipSuccess) return nullptr;
        return d;
    }
};

static int down(double *h, const double *d, size_t n)
{
    DG_HIP(hipMemcpy(h, d, n * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

static int check_dims(int nx, int ny, int nz)
{
    if (nx < 3 || ny < 3 || nz < 3 || (long long)nx * ny * nz > (1LL << 31) - 1) {
        set_error("invalid extents %d x %d x %d", nx, ny, nz);
        return -1;
    }
    return 0;
}

static hipStream_t scratch_stream() { return nullptr; }

}  // namespace dg

using namespace dg;

extern "C" {

int dg_status(void) { return g_status; }
const char *dg_last_error(void) { return g_err; }

int dg_init(const char *path) { return init_from_input(path, nullptr); }

int dg_finalize(void)
{
    Ctx &c = *ctx();
    const int rc = output_shutdown();
    if (g_ready) free_ctx(c);
    g_ready = false;
    return rc;
}
}  // extern "C"

namespace dg {

// A fresh resident state of nx x ny x nz (every field zero, the tuning
// defaults applied); the caller sets the scalars, then calls
// resident_scalars_changed().
int resident_setup(int nx, int ny, int nz)
{
    Ctx &c = *ctx();
    if (g_ready) {
        free_ctx(c);
        g_ready = false;
    }
    g_status = 0;
    g_err[0] = 0;
    if (check_dims(nx, ny, nz)) return -1;
    // m*2**k + 2 grids are accepted for the tolerance mode; the reference's
    // 2**ng + 2 check moves to its FMG schedule (op_mglin)
    if (alloc_levels(c, nx, ny, nz, false)) { free_ctx(c); return -1; }
    if (alloc_fields(c)) { free_ctx(c); return -1; }
    c.use_graphs = g_graphs_default;
    c.rb_split = g_split_default;
    c.rb_tail = g_tail_default;
    c.rb_tiles = g_tiles_default;
    c.rb_tile_nu = g_tile_nu_default;
    c.pcg_lookahead = g_lookahead_default;
    c.mg_omega = g_omega_default;
    c.mg_nuc = g_nuc_default;
    g_ready = true;
    return 0;
}

// what follows from the physical scalars: FP16 face coefficients are used
// when 1/rho of both fluids fits FP16's normal range with margin (rho1 = 1
// after nondim)
void resident_scalars_changed()
{
    Ctx &c = *ctx();
    const double rlo = c.s.rho2 < 1.0 ? c.s.rho2 : 1.0, rhi = c.s.rho2 > 1.0 ? c.s.rho2 : 1.0;
    c.coef16_ok = (rlo > 0.0 && 1.0 / rlo <= 3.0e4 && 1.0 / rhi >= 1.0e-3) ? 1 : 0;
}

bool resident_ready() { return g_ready; }
double *resident_field(int id) { return g_ready ? field_ptr(*ctx(), id) : nullptr; }

// dg_init: initialize (main.f:121-183) -- readinput of the 38-line
// inputdata.in, memoryaloc, nondim, then the droplet IC (smooth2) on the
// device or, with restart = 1, iniget's records from save.dat beside
// inputdata.in; bc_c, bc_v, calcrho.  The parsed input through in_out.
int init_from_input(const char *path, Input *in_out)
{
    Ctx &c = *ctx();
    if (g_ready) {
        free_ctx(c);
        g_ready = false;
    }
    g_status = 0;
    g_err[0] = 0;
    RunInfo &ri = run_info();
    ri.valid = false;
    Input in;
    if (read_input(path, in)) return -1;
    // iniget runs after nondim (main.f:139-174): the scales come from
    // inputdata.in, the solver settings and output cadences from the file
    Input run = in;
    int t0 = 0;
    std::vector<double> uvwcc[4];
    if (in.restart != 0) {
        const std::string p(path);
        const size_t slash = p.find_last_of('/');
        const std::string save = (slash == std::string::npos ? std::string() : p.substr(0, slash + 1)) + "save.dat";
        if (read_restart_file(save.c_str(), run, &t0, uvwcc)) return -1;
    }
    if (run.icase != 2) {
        set_error("icase=%d: only the droplet case (icase=2) is on this path", run.icase);
        return -1;
    }
    if (run.hswitch != 0 && run.hswitch != 1) {
        set_error("hswitch=%d: 0 (off) or 1 (CalcH3d interface heights)", run.hswitch);
        return -1;
    }
    if (run.surfacet != 2 && run.surfacet != 5 && run.surfacet != 6) {
        set_error("surfacet=%d: CSF (2), SGIP (5) and SGIP2 (6) are on this path; CSS, CSF-BKZ and Meier "
                  "(1, 3, 4) are not",
                  run.surfacet);
        return -1;
    }
    if (run.iturb < 0 || run.iturb > 4) {
        // kepsmodel STOPs with 'iturb problem' (main.f:1966-1968)
        set_error("iturb=%d: 0 laminar, 1 and 2 mixing length, 3 k-epsilon, 4 anisotropic k-epsilon", run.iturb);
        return -1;
    }
    if (resident_setup(in.nx, in.ny, in.nz)) return -1;
    nondim(in, c.s);
    if (in.restart != 0) {
        // the values iniget reads into COMMON that the time step uses as they are
        c.s.maxcfl = run.maxcfl;
        c.s.omega = run.omega;
        c.s.cycles = run.cycles; c.s.npre = run.npre; c.s.ntop = run.ntop; c.s.npost = run.npost;
        c.s.iturb = run.iturb; c.s.convectype = run.convectype; c.s.hswitch = run.hswitch;
        c.s.tprint = run.tprint;
    }
    resident_scalars_changed();
    // cmu = 0.09 (main.f:347-353); mut, turbst and the Reynolds stresses start at 0
    if (mg_fill(c.g.n(), c.stream, c.cmu, 0.09)) return -1;
    if ((run.hswitch == 1 || run.surfacet == 5 || run.surfacet == 6) && alloc_heights(c)) return -1;
    if (run.iturb == 4 && alloc_reynolds(c)) return -1;
    if (in.restart != 0) {
        // no smooth2 on a restart: y0 and z0 keep nondim's channel centre
        // (main.f:409-410), which eddyvisc reads
        c.s.y0 = (double)(in.ny - 2) * c.s.h / 2.0;
        c.s.z0 = (double)(in.nz - 2) * c.s.h / 2.0;
        const size_t bytes = c.g.n() * sizeof(double);
        double *dst[4] = {c.u, c.v, c.w, c.lev[c.ng - 1].cc};
        for (int q = 0; q < 4; q++) DG_HIP(hipMemcpy(dst[q], uvwcc[q].data(), bytes, hipMemcpyHostToDevice));
    } else if (op_smooth2(c)) {
        return -1;
    }
    if (op_bc_c(c, c.lev[c.ng - 1].cc) || op_bc_v(c) || op_calcrho(c)) return -1;
    DG_HIP(hipStreamSynchronize(c.stream));
    ri.in = run;
    ri.t0 = t0;
    ri.valid = true;
    if (in_out) *in_out = run;
    return 0;
}

RunInfo &run_info()
{
    static RunInfo r;
    return r;
}

}  // namespace dg

extern "C" {

int dg_vofmethod(const int *t)
{
    if (require_ready()) return -1;
    return op_vofmethod(*ctx(), *t);
}
int dg_kepsmodel(const int *t)
{
    if (require_ready()) return -1;
    return op_kepsmodel(*ctx(), *t);
}
int dg_eddyvisc(void)
{
    if (require_ready()) return -1;
    return op_eddyvisc(*ctx());
}
int dg_uv2p(void)
{
    if (require_ready()) return -1;
    Ctx &c = *ctx();
    if (alloc_reynolds(c)) return -1;
    return op_uv2p(c);
}
int dg_init_droplets(const int *n, const double *params)
{
    if (require_ready()) return -1;
    Ctx &c = *ctx();
    if (*n < 1 || *n > 4096) {
        set_error("dg_init_droplets: %d droplets (1 to 4096)", *n);
        return -1;
    }
    for (int d = 0; d < *n; d++)
        if (!(params[7 * d + 3] > 0.0) || !std::isfinite(params[7 * d + 0]) || !std::isfinite(params[7 * d + 1]) ||
            !std::isfinite(params[7 * d + 2])) {
            set_error("dg_init_droplets: droplet %d has radius %g at (%g, %g, %g)", d, params[7 * d + 3],
                      params[7 * d + 0], params[7 * d + 1], params[7 * d + 2]);
            return -1;
        }
    if (op_droplets(c, *n, params) || op_bc_v(c) || op_calcrho(c)) return -1;
    DG_HIP(hipStreamSynchronize(c.stream));
    return 0;
}
int dg_smooth2_droplet(double *params)
{
    if (require_ready()) return -1;
    const Ctx &c = *ctx();
    // smooth2's centre and velocities (main.f:5684-5690, 5739-5745)
    params[0] = (double)(c.g.nx - 2) * c.s.h / 4.0;
    params[1] = c.s.y0;
    params[2] = c.s.z0;
    params[3] = c.s.radius;
    params[4] = 0.0;
    params[5] = c.s.yvel1;
    params[6] = c.s.zvel1;
    return 0;
}
int dg_checktime(const int *t)
{
    if (require_ready()) return -1;
    return op_checktime(*ctx(), *t);
}
int dg_csf(void)
{
    if (require_ready()) return -1;
    return op_csf(*ctx());
}
int dg_calch3d(void)
{
    if (require_ready()) return -1;
    return op_calch3d(*ctx());
}
int dg_sgip(void)
{
    if (require_ready()) return -1;
    return op_sgip(*ctx());
}
int dg_sgip2(void)
{
    if (require_ready()) return -1;
    return op_sgip2(*ctx());
}
int dg_solvelpres(const int *t)
{
    if (require_ready()) return -1;
    return op_solvelpres(*ctx(), *t);
}
int dg_cfl(double *cfl)
{
    if (require_ready()) return -1;
    return op_cfl(*ctx(), cfl);
}
int dg_calcrho(void)
{
    if (require_ready()) return -1;
    return op_calcrho(*ctx());
}
int dg_mglin(void)
{
    if (require_ready()) return -1;
    return op_mglin(*ctx());
}

int dg_step(const int *t)
{
    if (require_ready()) return -1;
    Ctx &c = *ctx();
    if (op_vofmethod(c, *t, true)) return -1;
    if (c.s.hswitch == 1 && op_calch3d(c)) return -1;  // main.f:30
    if ((c.s.iturb == 1 || c.s.iturb == 2) && op_eddyvisc(c)) return -1;  // main.f:35
    if ((c.s.iturb == 3 || c.s.iturb == 4) && op_kepsmodel(c, *t)) return -1;  // main.f:39-40
    if (op_checktime(c, *t)) return -1;
    if (c.s.surfacet == 2 && op_csf(c, false)) return -1;
    if (c.s.surfacet == 5 && op_sgip(c, false)) return -1;
    if (c.s.surfacet == 6 && op_sgip2(c, false)) return -1;
    if (op_solvelpres(c, *t)) return -1;
    return 0;
}

int dg_run(const int *t0, const int *nsteps)
{
    const bool out = output_active();
    for (int t = *t0; t < *t0 + *nsteps; t++)
        if ((out && dg_printout(&t)) || dg_step(&t)) return -1;
    Ctx &c = *ctx();
    DG_HIP(hipStreamSynchronize(c.stream));
    return 0;
}

int dg_get_field(const int *id, double *host)
{
    if (require_ready()) return -1;
    Ctx &c = *ctx();
    double *d = field_ptr(c, *id);
    if (!d) { set_error("unknown field id %d", *id); return -1; }
    DG_HIP(hipStreamSynchronize(c.stream));
    DG_HIP(hipMemcpy(host, d, c.g.n() * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int dg_set_field(const int *id, const double *host)
{
    if (require_ready()) return -1;
    Ctx &c = *ctx();
    double *d = field_ptr(c, *id);
    if (!d) { set_error("unknown field id %d", *id); return -1; }
    DG_HIP(hipStreamSynchronize(c.stream));
    DG_HIP(hipMemcpy(d, host, c.g.n() * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

int dg_get_state(dg_state *out)
{
    if (require_ready()) return -1;
    Ctx &c = *ctx();
    DG_HIP(hipStreamSynchronize(c.stream));
    out->tau = c.s.tau; out->time = c.s.time; out->tauh = c.s.tauh;
    out->tauh2 = c.s.tauh2; out->tswap = c.s.tswap;
    out->npre = c.s.npre; out->ntop = c.s.ntop; out->npost = c.s.npost;
    out->nx = c.g.nx; out->ny = c.g.ny; out->nz = c.g.nz; out->ng = c.ng;
    out->mg_iters = c.mg_iters_last; out->mg_relres = c.mg_relres_last;
    out->cfl = c.cfl_last; out->epsdiv = c.epsdiv_last;
    return 0;
}

int dg_set_state(const dg_state *in)
{
    if (require_ready()) return -1;
    Ctx &c = *ctx();
    c.s.tau = in->tau; c.s.time = in->time; c.s.tauh = in->tauh; c.s.tauh2 = in->tauh2; c.s.tswap = in->tswap;
    c.s.npre = in->npre; c.s.ntop = in->ntop; c.s.npost = in->npost;
    return 0;
}

int dg_clip_cc(void)
{
    if (require_ready()) return -1;
    Ctx &c = *ctx();
    double *cc = c.lev[c.ng - 1].cc;
    return launch_linear(c.g.n(), c.stream, [=] __device__(size_t p) {
        double v = cc[p];
        cc[p] = v > 1.0 ? 1.0 : (v < 0.0 ? 0.0 : v);
    });
}

int dg_set_solver(const int *mode, const double *tol, const int *nu1, const int *nu2, const int *maxit,
                  const int *guess)
{
    if (require_ready()) return -1;
    Ctx &c = *ctx();
    if (*mode < 0 || *mode > 2 || *nu1 < 0 || *nu2 < 0 || *maxit < 1 || !(*tol > 0.0) || *guess < 0 ||
        *guess > 2) {
        set_error("dg_set_solver: invalid arguments");
        return -1;
    }
    c.mg_mode = *mode; c.mg_tol = *tol; c.mg_nu1 = *nu1; c.mg_nu2 = *nu2; c.mg_maxit = *maxit; c.mg_guess = *guess;
    return 0;
}

int dg_set_ordered_sums(const int *on)
{
    if (require_ready()) return -1;
    ctx()->ordered_sums = *on ? 1 : 0;
    return 0;
}

int dg_set_smoother(const double *omega, const int *nu_coarse)
{
    if (!(*omega > 0.0 && *omega < 2.0) || *nu_coarse < 0) {
        set_error("dg_set_smoother: 0 < omega < 2 and nu_coarse >= 0 required");
        return -1;
    }
    g_omega_default = *omega;
    g_nuc_default = *nu_coarse;
    Ctx &c = *ctx();
    c.mg_omega = g_omega_default;
    c.mg_nuc = g_nuc_default;
    return 0;
}

int dg_set_vof_guard(const int *on)
{
    if (require_ready()) return -1;
    ctx()->vof_guard = *on ? 1 : 0;
    return 0;
}

int dg_set_tuning(const int *graphs, const int *split, const int *tail)
{
    if (*split < 0 || *split > 3 || *tail < 0 || *tail > 2) {
        set_error("dg_set_tuning: split must be 0..3 and tail 0..2");
        return -1;
    }
    g_graphs_default = *graphs ? 1 : 0;
    g_split_default = *split;
    g_tail_default = *tail;
    Ctx &c = *ctx();
    c.use_graphs = g_graphs_default;
    c.rb_split = g_split_default;
    c.rb_tail = g_tail_default;
    c.rb_tiles = g_tiles_default;
    c.rb_tile_nu = g_tile_nu_default;
    return 0;
}

int dg_set_tile_smoother(const int *on, const int *nu)
{
    if (*on < 0 || *on > 2 || *nu < 0) {
        set_error("dg_set_tile_smoother: on must be 0, 1 or 2 and nu >= 0");
        return -1;
    }
    g_tiles_default = *on;
    g_tile_nu_default = *nu;
    ctx()->rb_tiles = g_tiles_default;
    ctx()->rb_tile_nu = g_tile_nu_default;
    return 0;
}

int dg_set_lookahead(const double *ratio)
{
    if (!(*ratio == *ratio)) {
        set_error("dg_set_lookahead: ratio is NaN");
        return -1;
    }
    g_lookahead_default = *ratio;
    ctx()->pcg_lookahead = g_lookahead_default;
    return 0;
}

// ---- host-pointer kernels ----------------------------------------------------

int dg_relax(double *u, const double *lengthsize, const double *rhs, const double *a, const double *b,
             const double *c, const double *ei, const int *ibc, const int *nx, const int *ny, const int *nz,
             const double *omega)
{
    (void)ibc;
    if (check_dims(*nx, *ny, *nz)) return -1;
    const Grid g = make_grid(*nx, *ny, *nz);
    const size_t n = g.n();
    DevBufs B;
    double *du = B.up(u, n), *dr = B.up(rhs, n), *da = B.up(a, n), *db = B.up(b, n), *dc = B.up(c, n),
           *de = B.up(ei, n);
    if (!du || !dr || !da || !db || !dc || !de) { set_error("dg_relax: device allocation failed"); return -1; }
    if (mg_relax(g, scratch_stream(), du, *lengthsize, dr, da, db, dc, de, *omega)) return -1;
    return down(u, du, n);
}

int dg_resid(const double *lengthsize, double *res, const double *u, const double *rhs, const double *a,
             const double *b, const double *c, const int *ibc, const int *nx, const int *ny, const int *nz)
{
    (void)ibc;
    if (check_dims(*nx, *ny, *nz)) return -1;
    const Grid g = make_grid(*nx, *ny, *nz);
    const size_t n = g.n();
    DevBufs B;
    double *dres = B.up(res, n), *du = B.up(u, n), *dr = B.up(rhs, n), *da = B.up(a, n), *db = B.up(b, n),
           *dc = B.up(c, n);
    if (!dres || !du || !dr || !da || !db || !dc) { set_error("dg_resid: device allocation failed"); return -1; }
    if (mg_resid(g, scratch_stream(), *lengthsize, dres, du, dr, da, db, dc)) return -1;
    return down(res, dres, n);
}

int dg_rstrct(double *uc, const double *uf, const int *ibc, const int *ncx, const int *ncy, const int *ncz)
{
    (void)ibc;
    if (check_dims(*ncx, *ncy, *ncz)) return -1;
    const Grid gc = make_grid(*ncx, *ncy, *ncz);
    const Grid gf = make_grid(2 * *ncx - 2, 2 * *ncy - 2, 2 * *ncz - 2);
    DevBufs B;
    double *duc = B.up(uc, gc.n()), *duf = B.up(uf, gf.n());
    if (!duc || !duf) { set_error("dg_rstrct: device allocation failed"); return -1; }
    if (mg_rstrct(gc, scratch_stream(), duc, duf)) return -1;
    return down(uc, duc, gc.n());
}

int dg_interp(double *uf, const double *uc, const int *ibc, const int *nfx, const int *nfy, const int *nfz)
{
    (void)ibc;
    if (check_dims(*nfx, *nfy, *nfz)) return -1;
    const Grid gf = make_grid(*nfx, *nfy, *nfz);
    const Grid gc = make_grid(*nfx / 2 + 1, *nfy / 2 + 1, *nfz / 2 + 1);
    DevBufs B;
    double *duf = B.up(uf, gf.n()), *duc = B.up(uc, gc.n());
    if (!duf || !duc) { set_error("dg_interp: device allocation failed"); return -1; }
    if (mg_interp(gf, scratch_stream(), duf, duc)) return -1;
    return down(uf, duf, gf.n());
}

int dg_addint(double *uf, const double *uc, double *res, const int *ibc, const int *nfx, const int *nfy,
              const int *nfz)
{
    (void)ibc;
    if (check_dims(*nfx, *nfy, *nfz)) return -1;
    const Grid gf = make_grid(*nfx, *nfy, *nfz);
    const Grid gc = make_grid(*nfx / 2 + 1, *nfy / 2 + 1, *nfz / 2 + 1);
    DevBufs B;
    double *duf = B.up(uf, gf.n()), *duc = B.up(uc, gc.n()), *dres = B.up(res, gf.n());
    if (!duf || !duc || !dres) { set_error("dg_addint: device allocation failed"); return -1; }
    if (mg_addint(gf, scratch_stream(), duf, duc, dres)) return -1;
    if (down(uf, duf, gf.n())) return -1;
    return down(res, dres, gf.n());
}

int dg_acfromac(double *a, double *b, double *c, double *ei, const double *a2, const double *b2, const double *c2,
                const int *nx, const int *ny, const int *nz)
{
    if (check_dims(*nx, *ny, *nz)) return -1;
    const Grid gc = make_grid(*nx, *ny, *nz);
    const Grid gf = make_grid(2 * *nx - 2, 2 * *ny - 2, 2 * *nz - 2);
    DevBufs B;
    double *da = B.up(a, gc.n()), *db = B.up(b, gc.n()), *dc = B.up(c, gc.n()), *de = B.up(ei, gc.n());
    double *da2 = B.up(a2, gf.n()), *db2 = B.up(b2, gf.n()), *dc2 = B.up(c2, gf.n());
    if (!da || !db || !dc || !de || !da2 || !db2 || !dc2) { set_error("dg_acfromac: device allocation failed"); return -1; }
    if (mg_acfromac(gc, scratch_stream(), da, db, dc, de, da2, db2, dc2)) return -1;
    if (down(a, da, gc.n()) || down(b, db, gc.n()) || down(c, dc, gc.n())) return -1;
    return down(ei, de, gc.n());
}

int dg_slvsml(double *u, const double *rhs, const double *a, const double *b, const double *c, const double *ei,
              const int *ibc, const int *nx, const int *ny, const int *nz, const int *ntop, const double *omega,
              const double *lengthsize)
{
    (void)ibc;
    if (check_dims(*nx, *ny, *nz)) return -1;
    const Grid g = make_grid(*nx, *ny, *nz);
    const size_t n = g.n();
    DevBufs B;
    double *du = B.up(u, n), *dr = B.up(rhs, n), *da = B.up(a, n), *db = B.up(b, n), *dc = B.up(c, n),
           *de = B.up(ei, n);
    if (!du || !dr || !da || !db || !dc || !de) { set_error("dg_slvsml: device allocation failed"); return -1; }
    for (int q = 1; q <= *ntop; q++)
        if (mg_relax(g, scratch_stream(), du, *lengthsize, dr, da, db, dc, de, *omega)) return -1;
    return down(u, du, n);
}

int dg_abcei(double *a, double *b, double *c, double *ei, const double *cc, const double *r1, const double *r2,
             const int *ibc, const int *nx, const int *ny, const int *nz)
{
    (void)ibc;
    if (check_dims(*nx, *ny, *nz)) return -1;
    const Grid g = make_grid(*nx, *ny, *nz);
    const size_t n = g.n();
    DevBufs B;
    double *da = B.up(a, n), *db = B.up(b, n), *dc = B.up(c, n), *de = B.up(ei, n), *dcc = B.up(cc, n);
    if (!da || !db || !dc || !de || !dcc) { set_error("dg_abcei: device allocation failed"); return -1; }
    if (mg_abcei(g, scratch_stream(), da, db, dc, de, dcc, *r1, *r2)) return -1;
    if (down(a, da, n) || down(b, db, n) || down(c, dc, n)) return -1;
    return down(ei, de, n);
}

int dg_bc_scalar(double *scal, const int *ibc, const int *nx, const int *ny, const int *nz)
{
    (void)ibc;
    if (check_dims(*nx, *ny, *nz)) return -1;
    const Grid g = make_grid(*nx, *ny, *nz);
    DevBufs B;
    double *d = B.up(scal, g.n());
    if (!d) { set_error("dg_bc_scalar: device allocation failed"); return -1; }
    if (op_bc_scalar(g, scratch_stream(), d)) return -1;
    return down(scal, d, g.n());
}

int dg_filterc27(const double *cc, double *cfilt, const int *nx, const int *ny, const int *nz)
{
    if (check_dims(*nx, *ny, *nz)) return -1;
    const Grid g = make_grid(*nx, *ny, *nz);
    DevBufs B;
    double *dcc = B.up(cc, g.n()), *dcf = B.up(cfilt, g.n());
    if (!dcc || !dcf) { set_error("dg_filterc27: device allocation failed"); return -1; }
    if (mg_filterc27(g, scratch_stream(), dcc, dcf)) return -1;
    return down(cfilt, dcf, g.n());
}

int dg_al3d_batch(const int *n, const double *b1, const double *b2, const double *b3, const double *cc, double *out)
{
    if (*n <= 0) return 0;
    DevBufs B;
    double *d1 = B.up(b1, *n), *d2 = B.up(b2, *n), *d3 = B.up(b3, *n), *dc = B.up(cc, *n), *dout = B.up(nullptr, *n);
    if (!d1 || !d2 || !d3 || !dc || !dout) { set_error("dg_al3d_batch: device allocation failed"); return -1; }
    if (plic_al3d_batch(*n, scratch_stream(), d1, d2, d3, dc, dout)) return -1;
    return down(out, dout, *n);
}

int dg_fl3d_batch(const int *n, const double *m1, const double *m2, const double *m3, const double *alpha,
                  const double *r0, const double *dr0, double *out)
{
    if (*n <= 0) return 0;
    DevBufs B;
    double *d1 = B.up(m1, *n), *d2 = B.up(m2, *n), *d3 = B.up(m3, *n), *da = B.up(alpha, *n), *dr = B.up(r0, *n),
           *dd = B.up(dr0, *n), *dout = B.up(nullptr, *n);
    if (!d1 || !d2 || !d3 || !da || !dr || !dd || !dout) { set_error("dg_fl3d_batch: device allocation failed"); return -1; }
    if (plic_fl3d_batch(*n, scratch_stream(), d1, d2, d3, da, dr, dd, dout)) return -1;
    return down(out, dout, *n);
}

int dg_poisson_solve(const int *nx, const int *ny, const int *nz, const double *lengthsize, const double *rhs,
                     const double *a, const double *b, const double *c, double *p, const int *mode, const double *tol,
                     const int *maxit, int *iters, double *relres)
{
    if (check_dims(*nx, *ny, *nz)) return -1;
    if (*mode < 1 || *mode > 2) {
        set_error("dg_poisson_solve: mode must be 1 (FP64) or 2 (FP32 V-cycle)");
        return -1;
    }
    Ctx lc;
    if (alloc_levels(lc, *nx, *ny, *nz, false)) { free_ctx(lc); return -1; }
    lc.pprev = nullptr;
    const size_t n = lc.g.n();
    if (dalloc(&lc.pprev, n)) { free_ctx(lc); return -1; }
    Level &F = lc.lev[lc.ng - 1];
    int rc = 0;
    rc |= hipMemcpy(F.rhs, rhs, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess;
    rc |= hipMemcpy(F.a, a, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess;
    rc |= hipMemcpy(F.b, b, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess;
    rc |= hipMemcpy(F.c, c, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess;
    rc |= hipMemcpy(lc.pprev, p, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess;
    if (rc) { set_error("dg_poisson_solve: upload failed"); free_ctx(lc); return -1; }
    lc.s.lengthsize = *lengthsize;
    lc.mg_mode = *mode; lc.mg_tol = *tol; lc.mg_maxit = *maxit; lc.mg_guess = 1;
    lc.use_graphs = g_graphs_default;
    lc.rb_split = g_split_default;
    lc.rb_tail = g_tail_default;
    lc.rb_tiles = g_tiles_default;
    lc.rb_tile_nu = g_tile_nu_default;
    lc.pcg_lookahead = g_lookahead_default;
    lc.coef16_ok = 0;  // arbitrary coefficients: FP32 in the split V-cycle
    lc.mg_omega = g_omega_default;
    lc.mg_nuc = g_nuc_default;
    if (op_mg_solve_tol(lc)) {
        // the iterations done and the residual reached stay readable
        *iters = lc.mg_iters_last;
        *relres = lc.mg_relres_last;
        free_ctx(lc);
        return -1;
    }
    if (hipStreamSynchronize(lc.stream) != hipSuccess ||
        hipMemcpy(p, F.u, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("dg_poisson_solve: download failed");
        free_ctx(lc);
        return -1;
    }
    *iters = lc.mg_iters_last;
    *relres = lc.mg_relres_last;
    free_ctx(lc);
    return 0;
}

int dg_profile_start(const int *kernel_id)
{
    Prof &p = prof();
    if (p.used) {
        Ctx &c = *ctx();
        if (c.stream) (void)hipStreamSynchronize(c.stream);
    }
    p.kernel = *kernel_id;
    p.used = 0;
    p.bytes = 0.0;
    p.launches = 0;
    return 0;
}

int dg_profile_stop(double *total_ms, int *launches, double *total_bytes)
{
    Prof &p = prof();
    const int id = p.kernel;
    p.kernel = PK_NONE;
    *total_ms = 0.0;
    *launches = 0;
    *total_bytes = 0.0;
    hipError_t e0 = hipDeviceSynchronize();
    if (e0 != hipSuccess) {
        set_error("dg_profile_stop(kernel %d): hipDeviceSynchronize: %s", id, hipGetErrorString(e0));
        p.used = 0;
        return -1;
    }
    double ms = 0.0;
    for (size_t q = 0; q + 1 < p.used; q += 2) {
        float e = 0.0f;
        hipError_t r = hipEventSynchronize(p.ev[q + 1]);
        if (r == hipSuccess) r = hipEventElapsedTime(&e, p.ev[q], p.ev[q + 1]);
        if (r != hipSuccess) {
            set_error("dg_profile_stop(kernel %d): event pair %zu of %zu: %s", id, q / 2, p.used / 2,
                      hipGetErrorString(r));
            p.used = 0;
            return -1;
        }
        ms += e;
    }
    *total_ms = ms;
    *launches = (int)p.launches;
    *total_bytes = p.bytes;
    p.used = 0;
    return 0;
}

}  // extern "C"
