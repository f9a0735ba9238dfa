// dg_fortran.hip -- the reference's own call forms: the Fortran-77 external
// procedures of the time step (implicit interfaces, every argument by
// reference, lowercase names with a trailing underscore), bound to the
// COMMON blocks of the Fortran program the library is linked into.
//
// A program built from main.f links against libdroplet_gpu.so with the
// time-step units taken out of main.f; its loop (main.f:19-73) then calls
// these definitions unchanged:
//   call VOFmethod(t)                                  -> vofmethod_
//   call kepsmodel(t, z(iu(ng)), z(ires(ng)))          -> kepsmodel_
//   call checktime(t)                                  -> checktime_
//   call csf(z(icc(ng)), z(iu(ng)), z(ires(ng)), z(irhs(ng)))        -> csf_
//   call solvelpres(z(icc(ng)), ..., z(iei(ng)), t)    -> solvelpres_
//
// State.  The library reads the program's COMMON blocks (input.h:19-95)
// through weak references to their symbols (memory_, pointers_, surfcom_,
// turbulence_, dimension_, diff_, cycling_, times_, dinem_, physics_); the
// array extents follow from /dimension/ (input.h's NMX, NMY, NMZ equal nx,
// ny, nz, readinput main.f:240-242), NGM and MEMLEN from the sizes of the
// pointers_ and memory_ symbols.  On the first call (or dg_attach_) every
// field and every multigrid heap slot is uploaded; from then on the device
// copies are authoritative and stay resident.  An array argument names a
// resident array by its address: z(icc(ng)) is the finest level's cc slot,
// u the /surfcom/ u, and so on; an address that is no resident array is an
// error.  Scalars are the host's: before each call the library reads tau,
// time, tauh, tauh2, tswap, NPRE, NTOP, NPOST from COMMON, and after it
// writes back those the routine changes (the 8-byte hand-backs).  Fields
// reach host memory through dg_sync_to_host_(mask), and at the end of
// solvelpres_ for step t the library downloads into COMMON the fields the
// program's own printout reads at the start of step t + 1 (main.f:21-23,
// 5020-5100), so the program's output files need no change to its loop.
//
// Errors: the reference STOPs on a fatal condition; these entry points
// print the library's error message and stop the program the same way
// (exit status 2), unless dg_set_stop_on_error_(0) asked for the error to
// be kept for dg_status_ instead.

#include <hip/hip_runtime.h>

#include <dlfcn.h>
#include <link.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "droplet_gpu.h"
#include "dg_ctx.h"
#include "dg_resident.h"

// COMMON blocks of the host program (absent: null)
extern "C" {
extern char memory_[] __attribute__((weak));
extern char pointers_[] __attribute__((weak));
extern char surfcom_[] __attribute__((weak));
extern char turbulence_[] __attribute__((weak));
extern char dimension_[] __attribute__((weak));
extern char diff_[] __attribute__((weak));
extern char cycling_[] __attribute__((weak));
extern char times_[] __attribute__((weak));
extern char dinem_[] __attribute__((weak));
extern char physics_[] __attribute__((weak));
extern char neww_[] __attribute__((weak));  // CalcH3d heights (input.h:90-95), optional
}

namespace dg {
namespace {

// heap slot kinds in the order of COMMON /pointers/ (input.h:19-21)
enum { P_CC = 0, P_U, P_RES, P_RHS, P_A, P_B, P_C, P_EI, P_KINDS };

// the host's COMMON blocks, resolved
struct HostCommon {
    bool bound = false;
    int nx = 0, ny = 0, nz = 0, NGM = 0, ng = 0;
    size_t N = 0;
    long memlen = 0;
    double *z = nullptr;
    int *mem = nullptr;
    int *ptrs = nullptr;  // icc(NGM), iu, ires, irhs, ia, ib, ic, iei, ng
    double *surf = nullptr, *turb = nullptr, *phys = nullptr, *diffd = nullptr, *cycd = nullptr, *dinem = nullptr;
    int *dim = nullptr, *tauh2 = nullptr, *cyci = nullptr, *times = nullptr;
    double *neww = nullptr;  // /NEWW/ hxmax, hxmin, hymax, hymin, hzmax, hzmin (null when absent)

    // /pointers/ entry of kind q (P_*) at Fortran level j (1-based)
    int &ptr(int q, int j) const { return ptrs[q * NGM + (j - 1)]; }
    double *slot(int q, int j) const { return z + (ptr(q, j) - 1); }
    // /surfcom/ u, v, w, S11, S22, S33, S12, S13, S23
    double *surfcom(int q) const { return surf + (size_t)q * N; }
    // /turbulence/ mut, lmix(NMX), turbst, ktur, eps, cmu, ...
    double *mut() const { return turb; }
    double *turbst() const { return turb + N + nx; }
    double *ktur() const { return turb + 2 * N + nx; }
    double *eps() const { return turb + 3 * N + nx; }
    double *cmu() const { return turb + 4 * N + nx; }
    // /PHYSICS/ 19 scalars, then rhon, mun
    double *rhon() const { return phys + 19; }
    double *mun() const { return phys + 19 + N; }
    // /CYCLING/ integers after MAXCFL, OMEGA
    enum { C_CYCLES = 0, C_NPRE, C_NPOST, C_NTOP, C_RESTART, C_ICASE, C_IJET, C_IDROP, C_IBUBBLE, C_ISHEET,
           C_ITURB, C_NFILTER, C_IBC, C_TSWAP, C_CONVECTYPE, C_HSWITCH, C_SURFACET };
};

HostCommon g_hc;
bool g_attached = false;
int g_stop_on_error = 1;

long symbol_size(const void *p)
{
    Dl_info info;
    const ElfW(Sym) *sym = nullptr;
    if (!p || !dladdr1(p, &info, (void **)&sym, RTLD_DL_SYMENT) || !sym) return -1;
    return (long)sym->st_size;
}

// memoryaloc's level count (main.f:287-297)
int reference_levels(int nx, int ny, int nz)
{
    int nn = nx < ny ? nx : ny;
    if (nz < nn) nn = nz;
    int j = nn - 2, ng = 1;
    while (j > 2) { j = j / 2; ng++; }
    return ng;
}

// level extents, finest first: n -> n/2 + 1 (memoryaloc)
void level_dims(int nx, int ny, int nz, int ng, int dims[][3])
{
    int a = nx, b = ny, c = nz;
    for (int j = ng; j >= 1; j--) {
        dims[j][0] = a; dims[j][1] = b; dims[j][2] = c;
        a = a / 2 + 1; b = b / 2 + 1; c = c / 2 + 1;
    }
}

int bind_common(HostCommon &h)
{
    if (!memory_ || !pointers_ || !surfcom_ || !turbulence_ || !dimension_ || !diff_ || !cycling_ || !times_ ||
        !dinem_ || !physics_) {
        set_error("the process has no Fortran COMMON blocks of input.h (memory_, pointers_, surfcom_, ...): the "
                  "Fortran call forms need a program built with input.h; C and Python hosts use dg_init");
        return -1;
    }
    h = HostCommon();
    h.dim = (int *)dimension_;
    h.nx = h.dim[0]; h.ny = h.dim[1]; h.nz = h.dim[2];
    if (h.nx < 3 || h.ny < 3 || h.nz < 3) {
        set_error("/dimension/ holds %d x %d x %d: call after the program's initialize (readinput)", h.nx, h.ny,
                  h.nz);
        return -1;
    }
    h.N = (size_t)h.nx * h.ny * h.nz;
    // NGM from the size of /pointers/ (8 arrays of NGM integers and ng)
    const long psz = symbol_size(pointers_);
    if (psz > 0) {
        h.NGM = (int)((psz / 4 - 1) / 8);
    } else {
        // symbol sizes unavailable: input.h's sizing, NGM the largest NGMX/NGMY/NGMZ
        int g = 0;
        const int nmax = h.nx > h.ny ? (h.nx > h.nz ? h.nx : h.nz) : (h.ny > h.nz ? h.ny : h.nz);
        while ((1 << g) + 2 < nmax) g++;
        h.NGM = g;
    }
    if (h.NGM < 1 || h.NGM > MAXLEV) {
        set_error("/pointers/ gives NGM = %d", h.NGM);
        return -1;
    }
    const long NN3 = ((1L << (3 * h.NGM)) - 1) / 7, NN2 = 1L << (2 * h.NGM), NN1 = 1L << h.NGM;
    const long msz = symbol_size(memory_);
    h.memlen = msz > 0 ? (msz - 4) / 8 : 64 * NN3 + 64 * NN2 + 192 * NN1 + 72 * h.NGM;
    const long ssz = symbol_size(surfcom_);
    if (ssz > 0 && (size_t)ssz != 9 * h.N * sizeof(double)) {
        set_error("/surfcom/ is %ld bytes, not 9 arrays of %d x %d x %d: input.h's NMX, NMY, NMZ differ from "
                  "nx, ny, nz (readinput, main.f:240-242)", ssz, h.nx, h.ny, h.nz);
        return -1;
    }
    h.z = (double *)memory_;
    h.mem = (int *)(memory_ + h.memlen * sizeof(double));
    h.ptrs = (int *)pointers_;
    h.surf = (double *)surfcom_;
    h.turb = (double *)turbulence_;
    h.phys = (double *)physics_;
    h.diffd = (double *)diff_;
    h.tauh2 = (int *)(diff_ + 4 * sizeof(double));
    h.cycd = (double *)cycling_;
    h.cyci = (int *)(cycling_ + 2 * sizeof(double));
    h.times = (int *)times_;
    h.dinem = (double *)dinem_;
    h.neww = neww_ ? (double *)neww_ : nullptr;
    h.bound = true;
    return 0;
}

// every heap slot z(i*(j)) is preceded by MYMALLOC's length word
// (main.f:4541-4554); checks the program's /pointers/ against the level
// extents
int check_heap(const HostCommon &h)
{
    const int ng = h.ptrs[8 * h.NGM];
    if (ng != reference_levels(h.nx, h.ny, h.nz)) {
        set_error("/pointers/ ng = %d, expected %d for %d x %d x %d: call after the program's memoryaloc", ng,
                  reference_levels(h.nx, h.ny, h.nz), h.nx, h.ny, h.nz);
        return -1;
    }
    int dims[MAXLEV + 1][3];
    level_dims(h.nx, h.ny, h.nz, ng, dims);
    for (int j = 1; j <= ng; j++) {
        const long len = (long)dims[j][0] * dims[j][1] * dims[j][2];
        for (int q = 0; q < P_KINDS; q++) {
            const long p = h.ptr(q, j);
            if (p < 2 || p - 1 + len > h.memlen || h.z[p - 2] != (double)len) {
                set_error("heap slot %d of level %d (z index %ld) does not hold MYMALLOC's length %ld", q, j, p, len);
                return -1;
            }
        }
    }
    return 0;
}

Level &finest() { return ctx()->lev[ctx()->ng - 1]; }

// device array of heap kind q at Fortran level j
double *dev_slot(int q, int j)
{
    Level &L = ctx()->lev[j - 1];
    double *const s[P_KINDS] = {L.cc, L.u, L.res, L.rhs, L.a, L.b, L.c, L.ei};
    return s[q];
}

// the host arrays a field id (DG_F_*) lives in, nullptr for none
double *host_field(int id)
{
    const HostCommon &h = g_hc;
    const int ng = h.ng;
    switch (id) {
    case DG_F_CC: return h.slot(P_CC, ng);
    case DG_F_P: return h.slot(P_U, ng);
    case DG_F_U: case DG_F_V: case DG_F_W: return h.surfcom(id - DG_F_U);
    case DG_F_S11: case DG_F_S22: case DG_F_S33: case DG_F_S12: case DG_F_S13: case DG_F_S23:
        return h.surfcom(3 + id - DG_F_S11);
    case DG_F_RHON: return h.rhon();
    case DG_F_MUN: return h.mun();
    case DG_F_MUT: return h.mut();
    case DG_F_KTUR: return h.ktur();
    case DG_F_EPS: return h.eps();
    case DG_F_A: return h.slot(P_A, ng);
    case DG_F_B: return h.slot(P_B, ng);
    case DG_F_C: return h.slot(P_C, ng);
    case DG_F_EI: return h.slot(P_EI, ng);
    case DG_F_RHS: return h.slot(P_RHS, ng);
    case DG_F_RES: return h.slot(P_RES, ng);
    // /NEWW/ holds hxmax, hxmin, hymax, hymin, hzmax, hzmin in that order
    case DG_F_HXMIN: return h.neww ? h.neww + 1 * h.N : nullptr;
    case DG_F_HXMAX: return h.neww ? h.neww + 0 * h.N : nullptr;
    case DG_F_HYMIN: return h.neww ? h.neww + 3 * h.N : nullptr;
    case DG_F_HYMAX: return h.neww ? h.neww + 2 * h.N : nullptr;
    case DG_F_HZMIN: return h.neww ? h.neww + 5 * h.N : nullptr;
    case DG_F_HZMAX: return h.neww ? h.neww + 4 * h.N : nullptr;
    // /TURBULENCE/ u2p, v2p, w2p, uvp, uwp, vwp follow cmu (input.h:36-43)
    case DG_F_U2P: case DG_F_V2P: case DG_F_W2P: case DG_F_UVP: case DG_F_UWP: case DG_F_VWP:
        return h.cmu() + (size_t)(id - DG_F_U2P + 1) * h.N;
    case DG_F_CMU: return h.cmu();
    default: return nullptr;
    }
}

// The resident array a host address names (the start of a finest-level
// heap slot or of a COMMON field), nullptr for none.
double *resident_of(const double *host)
{
    if (!host) return nullptr;
    for (int id = 0; id < DG_NFIELDS; id++) {
        const double *hp = host_field(id);
        if (hp && hp == host) return resident_field(id);
    }
    return nullptr;
}

int expect_slot(const char *who, int argno, const double *host, const double *dev_expected, const char *what)
{
    const double *d = resident_of(host);
    if (d != dev_expected) {
        set_error("%s: argument %d is not %s (the reference's call form, main.f:23-69)", who, argno, what);
        return -1;
    }
    return 0;
}

// scalars the time-step routines read and change: COMMON -> resident
void pull_scalars()
{
    Scalars &s = ctx()->s;
    const HostCommon &h = g_hc;
    s.tau = h.phys[0];
    s.time = h.phys[1];
    s.tauh = h.diffd[3];
    s.tauh2 = *h.tauh2;
    s.tswap = h.cyci[HostCommon::C_TSWAP];
    s.npre = h.cyci[HostCommon::C_NPRE];
    s.npost = h.cyci[HostCommon::C_NPOST];
    s.ntop = h.cyci[HostCommon::C_NTOP];
}

// and back (the 8-byte hand-backs)
void push_scalars()
{
    const Scalars &s = ctx()->s;
    const HostCommon &h = g_hc;
    h.phys[0] = s.tau;
    h.phys[1] = s.time;
    h.diffd[3] = s.tauh;
    *h.tauh2 = s.tauh2;
    h.cyci[HostCommon::C_TSWAP] = s.tswap;
    h.cyci[HostCommon::C_NPRE] = s.npre;
    h.cyci[HostCommon::C_NPOST] = s.npost;
    h.cyci[HostCommon::C_NTOP] = s.ntop;
}

// every scalar of the path: COMMON -> resident (after the program's nondim)
void read_all_scalars()
{
    Scalars &s = ctx()->s;
    const HostCommon &h = g_hc;
    const double *p = h.phys;
    s.tau = p[0]; s.time = p[1]; s.radius = p[2]; s.h = p[3]; s.rho1 = p[4]; s.rho2 = p[5]; s.mu1 = p[6];
    s.mu2 = p[7]; s.sigma = p[8]; s.g = p[9]; s.g1 = p[10]; s.y0 = p[11]; s.z0 = p[12];
    s.xvel1 = p[13]; s.yvel1 = p[14]; s.zvel1 = p[15]; s.xvel2 = p[16]; s.yvel2 = p[17]; s.zvel2 = p[18];
    s.tauh = h.diffd[3];
    s.tauh2 = *h.tauh2;
    s.maxcfl = h.cycd[0];
    s.omega = h.cycd[1];
    const int *ci = h.cyci;
    s.cycles = ci[HostCommon::C_CYCLES]; s.npre = ci[HostCommon::C_NPRE]; s.npost = ci[HostCommon::C_NPOST];
    s.ntop = ci[HostCommon::C_NTOP]; s.iturb = ci[HostCommon::C_ITURB]; s.tswap = ci[HostCommon::C_TSWAP];
    s.convectype = ci[HostCommon::C_CONVECTYPE]; s.hswitch = ci[HostCommon::C_HSWITCH];
    s.surfacet = ci[HostCommon::C_SURFACET];
    s.tprint = h.times[5];
    // /dinem/ tau_SI, timescale, lengthscale, length_SI, ..., lengthsize (11th)
    s.tau_si = h.dinem[0]; s.timescale = h.dinem[1]; s.lengthscale = h.dinem[2]; s.lengthsize = h.dinem[10];
}

// every scalar: resident -> COMMON, and the input's counters (dg_init_input_)
void write_all_scalars(const Input &in)
{
    const Scalars &s = ctx()->s;
    const HostCommon &h = g_hc;
    double *p = h.phys;
    p[0] = s.tau; p[1] = s.time; p[2] = s.radius; p[3] = s.h; p[4] = s.rho1; p[5] = s.rho2; p[6] = s.mu1;
    p[7] = s.mu2; p[8] = s.sigma; p[9] = s.g; p[10] = s.g1; p[11] = s.y0; p[12] = s.z0;
    p[13] = s.xvel1; p[14] = s.yvel1; p[15] = s.zvel1; p[16] = s.xvel2; p[17] = s.yvel2; p[18] = s.zvel2;
    h.diffd[3] = s.tauh;
    *h.tauh2 = s.tauh2;
    h.cycd[0] = s.maxcfl;
    h.cycd[1] = s.omega;
    int *ci = h.cyci;
    ci[HostCommon::C_CYCLES] = s.cycles; ci[HostCommon::C_NPRE] = s.npre; ci[HostCommon::C_NPOST] = s.npost;
    ci[HostCommon::C_NTOP] = s.ntop; ci[HostCommon::C_RESTART] = in.restart; ci[HostCommon::C_ICASE] = in.icase;
    ci[HostCommon::C_IDROP] = in.icase == 2 ? 1 : 0; ci[HostCommon::C_ITURB] = s.iturb;
    ci[HostCommon::C_NFILTER] = in.nfilter; ci[HostCommon::C_IBC] = in.ibc; ci[HostCommon::C_TSWAP] = s.tswap;
    ci[HostCommon::C_CONVECTYPE] = s.convectype; ci[HostCommon::C_HSWITCH] = s.hswitch;
    ci[HostCommon::C_SURFACET] = s.surfacet;
    // /TIMES/ t0, tmax, tstore, ntecplot, tsave, tprint
    h.times[0] = run_info().t0; h.times[1] = in.tmax; h.times[3] = in.ntecplot; h.times[4] = in.tsave; h.times[5] = in.tprint;
    double *d = h.dinem;
    d[0] = s.tau_si; d[1] = s.timescale; d[2] = s.lengthscale; d[3] = in.length; d[4] = in.rho1; d[5] = in.rho2;
    d[6] = in.visc1; d[7] = in.visc2; d[8] = in.sigma; d[9] = in.grav; d[10] = s.lengthsize; d[12] = in.grav1;
    d[13] = in.xvel1; d[14] = in.yvel1; d[15] = in.zvel1; d[16] = in.xvel2; d[17] = in.yvel2; d[18] = in.zvel2;
    d[19] = in.radius; d[20] = in.z0; d[21] = in.y0;
}

// memoryaloc's bookkeeping (main.f:274-322, MYMALLOC main.f:4541-4554) in the
// host's /pointers/ and /memory/
int write_heap_pointers(HostCommon &h)
{
    const int ng = reference_levels(h.nx, h.ny, h.nz);
    if (ng > h.NGM) {
        set_error("%d multigrid levels do not fit NGM = %d", ng, h.NGM);
        return -1;
    }
    int dims[MAXLEV + 1][3];
    level_dims(h.nx, h.ny, h.nz, ng, dims);
    long mem = 0;
    for (int j = ng; j >= 1; j--) {
        const long len = (long)dims[j][0] * dims[j][1] * dims[j][2];
        const int order[P_KINDS] = {P_RHS, P_RES, P_U, P_CC, P_A, P_B, P_C, P_EI};
        for (int q : order) {
            if (mem + len + 1 > h.memlen) {
                set_error("MEMLEN=%ld, mem+len+1: %ld: insufficient memory in MYMALLOC", h.memlen, mem + len + 1);
                return -1;
            }
            h.z[mem] = (double)len;
            h.ptr(q, j) = (int)(mem + 2);
            mem += len + 1;
        }
    }
    *h.mem = (int)mem;
    h.ptrs[8 * h.NGM] = ng;
    return 0;
}

// host <-> device copies of the heap slots of every level
int copy_heap(bool to_device)
{
    Ctx &c = *ctx();
    for (int j = 1; j <= g_hc.ng; j++) {
        const size_t n = c.lev[j - 1].g.n();
        for (int q = 0; q < P_KINDS; q++) {
            double *d = dev_slot(q, j), *hp = g_hc.slot(q, j);
            if (to_device) DG_HIP(hipMemcpy(d, hp, n * sizeof(double), hipMemcpyHostToDevice));
            else DG_HIP(hipMemcpy(hp, d, n * sizeof(double), hipMemcpyDeviceToHost));
        }
    }
    return 0;
}

// bit of a field id in the sync masks: ids below DG_F_U2P their own bit,
// cmu bit 29, the six Reynolds-stress arrays bit 31 (bit 30: the heap)
unsigned sync_bit(int id)
{
    if (id < DG_F_U2P) return 1u << id;
    return id == DG_F_CMU ? 1u << 29 : 1u << 31;
}

int copy_fields(unsigned mask, bool to_device)
{
    Ctx &c = *ctx();
    DG_HIP(hipStreamSynchronize(c.stream));
    const size_t bytes = g_hc.N * sizeof(double);
    for (int id = 0; id < DG_NFIELDS; id++) {
        if (!(mask & sync_bit(id))) continue;
        double *hp = host_field(id), *d = resident_field(id);
        if (!hp || !d) continue;
        if (to_device) DG_HIP(hipMemcpy(d, hp, bytes, hipMemcpyHostToDevice));
        else DG_HIP(hipMemcpy(hp, d, bytes, hipMemcpyDeviceToHost));
    }
    return 0;
}

// The fields the program's own printout (main.f:5020-5100) reads at the
// start of step t, into COMMON: u, v, w, cc and mut every tprint steps
// (velocityc, FINDNINF, SUMFIELD, FINDMAX, FINDN2), u, v, w, cc every
// ntecplot steps (printtecp) and at t = tsave (savedata), the pressure
// every 100 steps.  /TIMES/ holds ntecplot, tsave, tprint (input.h:67).
int printout_fields_to_host(int t)
{
    const int ntecplot = g_hc.times[3], tsave = g_hc.times[4], tprint = g_hc.times[5];
    const unsigned uvwc = (1u << DG_F_U) | (1u << DG_F_V) | (1u << DG_F_W) | (1u << DG_F_CC);
    unsigned m = 0;
    if (tprint > 0 && t % tprint == 0) m |= uvwc | (1u << DG_F_MUT);
    if (ntecplot > 0 && t % ntecplot == 0) m |= uvwc;
    if (t % 100 == 0) m |= 1u << DG_F_P;
    if (t == tsave) m |= uvwc;
    return m ? copy_fields(m, false) : 0;
}

constexpr unsigned HEAP_BIT = 1u << 30;        // every level's heap slots
constexpr unsigned ALL_FIELDS = ~HEAP_BIT;     // every field id, cmu and the Reynolds stresses
static_assert(DG_F_U2P == 29, "sync-mask bits 0-28 are the field ids below the Reynolds stresses");

int attach()
{
    if (bind_common(g_hc) || check_heap(g_hc)) return -1;
    g_hc.ng = g_hc.ptrs[8 * g_hc.NGM];
    if (resident_setup(g_hc.nx, g_hc.ny, g_hc.nz)) return -1;
    if (ctx()->ng != g_hc.ng) {
        set_error("%d x %d x %d: the program's memoryaloc has %d levels, the library %d", g_hc.nx, g_hc.ny, g_hc.nz,
                  g_hc.ng, ctx()->ng);
        return -1;
    }
    read_all_scalars();
    resident_scalars_changed();
    if ((ctx()->s.hswitch == 1 || ctx()->s.surfacet == 5 || ctx()->s.surfacet == 6) && alloc_heights(*ctx()))
        return -1;
    if (ctx()->s.iturb == 4 && alloc_reynolds(*ctx())) return -1;
    if (copy_heap(true) || copy_fields(ALL_FIELDS, true)) return -1;
    g_attached = true;
    return 0;
}

// entry of every resident call form: attach on first use, take the host's scalars
bool enter()
{
    if (!g_attached || !resident_ready()) {
        g_attached = false;
        if (attach()) return false;
    }
    pull_scalars();
    return true;
}

void fail(const char *who)
{
    fprintf(stderr, "%s: %s\n", who, last_error());
    fflush(stderr);
    if (g_stop_on_error) {
        fflush(stdout);
        exit(2);
    }
}

// runs one resident routine: attach, scalars in, body, scalars out
template <class F>
void call_form(const char *who, F body)
{
    if (!enter()) { fail(who); return; }
    if (body(*ctx())) { fail(who); return; }
    push_scalars();
}

}  // namespace
}  // namespace dg

using namespace dg;

extern "C" {

// ---- lifecycle -------------------------------------------------------------

// After the program's initialize (main.f:15): binds its COMMON blocks and
// uploads every field and heap slot.  The first call form does the same
// when this is not called.
void dg_attach_(void)
{
    g_attached = false;
    if (attach()) fail("dg_attach");
}

// For a program without the reference's initialize: readinput of
// ./inputdata.in, memoryaloc, nondim and the droplet IC run in the library
// (as dg_init), and their results are written to the program's COMMON
// blocks (scalars, /pointers/ and the heap bookkeeping, every field).
void dg_init_input_(void)
{
    Input in;
    g_attached = false;
    if (!dimension_) {
        bind_common(g_hc);  // names what is missing
        fail("dg_init_input");
        return;
    }
    if (init_from_input("inputdata.in", &in)) { fail("dg_init_input"); return; }
    // readinput's /dimension/ (main.f:188-190)
    int *dim = (int *)dimension_;
    dim[0] = in.nx; dim[1] = in.ny; dim[2] = in.nz;
    if (bind_common(g_hc) || write_heap_pointers(g_hc)) { fail("dg_init_input"); return; }
    g_hc.ng = g_hc.ptrs[8 * g_hc.NGM];
    write_all_scalars(in);
    if (copy_heap(false) || copy_fields(ALL_FIELDS, false)) { fail("dg_init_input"); return; }
    g_attached = true;
}

// Fields to host memory before the program's own readers.  mask: bit id
// for each DG_F_* field id below DG_F_U2P with a host array, bit 29 cmu,
// bit 31 the six Reynolds stresses, bit 30 every level's heap slots; -1
// all of them.
void dg_sync_to_host_(const int *mask)
{
    if (!enter()) { fail("dg_sync_to_host"); return; }
    const unsigned m = (unsigned)*mask;
    if (copy_fields(m & ALL_FIELDS, false) || ((m & HEAP_BIT) && copy_heap(false))) fail("dg_sync_to_host");
}

// Fields the program changed in host memory back to the device (same mask).
void dg_sync_to_device_(const int *mask)
{
    if (!enter()) { fail("dg_sync_to_device"); return; }
    const unsigned m = (unsigned)*mask;
    if (((m & HEAP_BIT) && copy_heap(true)) || copy_fields(m & ALL_FIELDS, true)) fail("dg_sync_to_device");
}

int dg_status_(void) { return dg_status(); }

void dg_set_stop_on_error_(const int *on) { g_stop_on_error = *on ? 1 : 0; }

void dg_finalize_(void)
{
    g_attached = false;
    dg_finalize();
}

// ---- the time step's routines in the reference's call forms -------------------

// VOFmethod(t), main.f:795-824
void vofmethod_(const int *t)
{
    call_form("VOFmethod", [&](Ctx &c) { return op_vofmethod(c, *t); });
}

// kepsmodel(t, kturnew, epsnew), main.f:1888-2026; called with z(iu(ng)), z(ires(ng))
void kepsmodel_(const int *t, double *kturnew, double *epsnew)
{
    call_form("kepsmodel", [&](Ctx &c) {
        if (expect_slot("kepsmodel", 2, kturnew, finest().u, "z(iu(ng))") ||
            expect_slot("kepsmodel", 3, epsnew, finest().res, "z(ires(ng))"))
            return -1;
        return op_kepsmodel(c, *t);
    });
}

// checktime(t), main.f:2156-2175
void checktime_(const int *t)
{
    call_form("checktime", [&](Ctx &c) { return op_checktime(c, *t); });
}

// DOUBLE PRECISION FUNCTION CFL(), main.f:4848-4881: max |u|,|v|,|w| * tauh
double cfl_(void)
{
    double v = 0.0;
    call_form("CFL", [&](Ctx &c) { return op_cfl(c, &v); });
    return v;
}

// csf(ccc, w11, w22, w33), main.f:2363-2481; called with z(icc(ng)),
// z(iu(ng)), z(ires(ng)), z(irhs(ng))
void csf_(double *ccc, double *w11, double *w22, double *w33)
{
    call_form("csf", [&](Ctx &c) {
        Level &L = finest();
        if (expect_slot("csf", 1, ccc, L.cc, "z(icc(ng))") || expect_slot("csf", 2, w11, L.u, "z(iu(ng))") ||
            expect_slot("csf", 3, w22, L.res, "z(ires(ng))") || expect_slot("csf", 4, w33, L.rhs, "z(irhs(ng))"))
            return -1;
        return op_csf(c);
    });
}

// solvelpres(cc, w1, w2, w3, a, b, c, ei, t), main.f:3236-3439
void solvelpres_(double *cc, double *w1, double *w2, double *w3, double *a, double *b, double *c, double *ei,
                 const int *t)
{
    call_form("solvelpres", [&](Ctx &cx) {
        Level &L = finest();
        if (expect_slot("solvelpres", 1, cc, L.cc, "z(icc(ng))") ||
            expect_slot("solvelpres", 2, w1, L.u, "z(iu(ng))") ||
            expect_slot("solvelpres", 3, w2, L.res, "z(ires(ng))") ||
            expect_slot("solvelpres", 4, w3, L.rhs, "z(irhs(ng))") ||
            expect_slot("solvelpres", 5, a, L.a, "z(ia(ng))") || expect_slot("solvelpres", 6, b, L.b, "z(ib(ng))") ||
            expect_slot("solvelpres", 7, c, L.c, "z(ic(ng))") || expect_slot("solvelpres", 8, ei, L.ei, "z(iei(ng))"))
            return -1;
        if (op_solvelpres(cx, *t)) return -1;
        return printout_fields_to_host(*t + 1);
    });
}

// eddyvisc(cc, ufilt), main.f:1828-1885; called with z(icc(ng)), z(iu(ng))
// (main.f:35): the filtered velocity goes into the pressure slot
void eddyvisc_(double *cc, double *ufilt)
{
    call_form("eddyvisc", [&](Ctx &c) {
        Level &L = finest();
        if (expect_slot("eddyvisc", 1, cc, L.cc, "z(icc(ng))") ||
            expect_slot("eddyvisc", 2, ufilt, L.u, "z(iu(ng))"))
            return -1;
        return op_eddyvisc(c);
    });
}

// uv2p(), main.f:2032-2155 (called by kepsmodel with iturb = 4, main.f:2023)
void uv2p_(void)
{
    call_form("uv2p", [&](Ctx &c) { return op_uv2p(c); });
}

// CalcH3d(cc), main.f:1312-1810 (called with z(icc(ng)), main.f:30)
void calch3d_(double *cc)
{
    call_form("CalcH3d", [&](Ctx &c) {
        if (expect_slot("CalcH3d", 1, cc, finest().cc, "z(icc(ng))")) return -1;
        if (alloc_heights(c)) return -1;
        return op_calch3d(c);
    });
}

// SGIP(ccc, w11, w22, w33), main.f:2835-3027, and SGIP2, main.f:3029-3166
// (called as csf, main.f:59-62)
static void sgip_form(const char *who, double *ccc, double *w11, double *w22, double *w33,
                      int (*op)(Ctx &, bool))
{
    call_form(who, [&](Ctx &c) {
        Level &L = finest();
        if (expect_slot(who, 1, ccc, L.cc, "z(icc(ng))") || expect_slot(who, 2, w11, L.u, "z(iu(ng))") ||
            expect_slot(who, 3, w22, L.res, "z(ires(ng))") || expect_slot(who, 4, w33, L.rhs, "z(irhs(ng))"))
            return -1;
        if (alloc_heights(c)) return -1;
        return op(c, true);
    });
}
void sgip_(double *ccc, double *w11, double *w22, double *w33)
{
    sgip_form("SGIP", ccc, w11, w22, w33, op_sgip);
}
void sgip2_(double *ccc, double *w11, double *w22, double *w33)
{
    sgip_form("SGIP2", ccc, w11, w22, w33, op_sgip2);
}

// calcrho(cc), main.f:2177-2190
void calcrho_(double *cc)
{
    call_form("calcrho", [&](Ctx &c) {
        if (expect_slot("calcrho", 1, cc, finest().cc, "z(icc(ng))")) return -1;
        return op_calcrho(c);
    });
}

// bc_c(cc), main.f:4331-4396, on the resident volume fraction
void bc_c_(double *cc)
{
    call_form("bc_c", [&](Ctx &c) {
        if (expect_slot("bc_c", 1, cc, finest().cc, "z(icc(ng))")) return -1;
        return op_bc_c(c, finest().cc);
    });
}

// mglin(r1, r2), main.f:3645-3763: the FMG schedule on the resident heap
// (z(irhs(ng)), z(ia(ng)), ... -> z(iu(ng))); r1 and r2 are unused there too
void mglin_(const double *r1, const double *r2)
{
    (void)r1;
    (void)r2;
    call_form("mglin", [&](Ctx &c) { return op_mglin(c); });
}

// swp(us, cc, vof1, vof2, vof3, d), main.f:848-1000: one split sweep along
// d with velocity us (u, v or w of /surfcom/) on resident arrays
void swp_(double *us, double *cc, double *vof1, double *vof2, double *vof3, const int *d)
{
    call_form("swp", [&](Ctx &c) {
        double *dus = resident_of(us), *dcc = resident_of(cc);
        double *d1 = resident_of(vof1), *d2 = resident_of(vof2), *d3 = resident_of(vof3);
        if (dus != c.u && dus != c.v && dus != c.w) {
            set_error("swp: argument 1 is not u, v or w of /surfcom/");
            return -1;
        }
        if (!dcc || !d1 || !d2 || !d3 || dcc == d1 || dcc == d2 || dcc == d3 || d1 == d2 || d1 == d3 || d2 == d3) {
            set_error("swp: arguments 2-5 must be four different resident arrays (z(icc(ng)), z(iei(ng)), "
                      "z(ires(ng)), z(irhs(ng)) in VOFmethod, main.f:806-816)");
            return -1;
        }
        return op_swp(c, dus, dcc, d1, d2, d3, *d);
    });
}

// convective(w1, w2, w3), main.f:3441-3580: u, v, w of /surfcom/ from the
// predicted velocities w1..w3 (resident arrays)
void convective_(double *w1, double *w2, double *w3)
{
    call_form("convective", [&](Ctx &c) {
        const double *d1 = resident_of(w1), *d2 = resident_of(w2), *d3 = resident_of(w3);
        if (!d1 || !d2 || !d3) {
            set_error("convective: arguments must be resident arrays (z(iu(ng)), z(ires(ng)), z(irhs(ng)) in "
                      "solvelpres, main.f:3338)");
            return -1;
        }
        if (d1 == c.u || d1 == c.v || d1 == c.w || d2 == c.u || d2 == c.v || d2 == c.w || d3 == c.u ||
            d3 == c.v || d3 == c.w) {
            set_error("convective: w1..w3 must not be u, v, w (they are its output)");
            return -1;
        }
        return op_convective(c, d1, d2, d3);
    });
}

// ---- fine-grained routines under their own names (host arrays) -------------
// The multigrid and PLIC kernels with the reference's argument lists, on
// whatever host arrays they are given: upload, run on the GPU, download
// (the dg_* entry points of the same lists).  For unit drop-in and tests;
// they need no COMMON blocks.

#define DG_FORM(who, call) \
    do {                   \
        if (call) fail(who); \
    } while (0)

// relax(u, lengthsize, rhs, a, b, c, ei, ibc, nx, ny, nz, OMEGA), main.f:3956
void relax_(double *u, const double *lengthsize, const double *rhs, const double *a, const double *b, const double *c,
            const double *ei, const int *ibc, const int *nx, const int *ny, const int *nz, const double *omega)
{
    DG_FORM("relax", dg_relax(u, lengthsize, rhs, a, b, c, ei, ibc, nx, ny, nz, omega));
}
// resid(lengthsize, res, u, rhs, a, b, c, ibc, nx, ny, nz), main.f:3902
void resid_(const double *lengthsize, double *res, const double *u, const double *rhs, const double *a,
            const double *b, const double *c, const int *ibc, const int *nx, const int *ny, const int *nz)
{
    DG_FORM("resid", dg_resid(lengthsize, res, u, rhs, a, b, c, ibc, nx, ny, nz));
}
// rstrct(uc, uf, ibc, ncx, ncy, ncz), main.f:3765
void rstrct_(double *uc, const double *uf, const int *ibc, const int *ncx, const int *ncy, const int *ncz)
{
    DG_FORM("rstrct", dg_rstrct(uc, uf, ibc, ncx, ncy, ncz));
}
// interp(uf, uc, ibc, nfx, nfy, nfz), main.f:3866
void interp_(double *uf, const double *uc, const int *ibc, const int *nfx, const int *nfy, const int *nfz)
{
    DG_FORM("interp", dg_interp(uf, uc, ibc, nfx, nfy, nfz));
}
// addint(uf, uc, res, ibc, nfx, nfy, nfz), main.f:3934
void addint_(double *uf, const double *uc, double *res, const int *ibc, const int *nfx, const int *nfy,
             const int *nfz)
{
    DG_FORM("addint", dg_addint(uf, uc, res, ibc, nfx, nfy, nfz));
}
// acfromac(a, b, c, ei, a2, b2, c2, nx, ny, nz), main.f:3793
void acfromac_(double *a, double *b, double *c, double *ei, const double *a2, const double *b2, const double *c2,
               const int *nx, const int *ny, const int *nz)
{
    DG_FORM("acfromac", dg_acfromac(a, b, c, ei, a2, b2, c2, nx, ny, nz));
}
// slvsml(u, rhs, a, b, c, ei, ibc, nx, ny, nz, NTOP, OMEGA, lengthsize), main.f:3846
void slvsml_(double *u, const double *rhs, const double *a, const double *b, const double *c, const double *ei,
             const int *ibc, const int *nx, const int *ny, const int *nz, const int *ntop, const double *omega,
             const double *lengthsize)
{
    DG_FORM("slvsml", dg_slvsml(u, rhs, a, b, c, ei, ibc, nx, ny, nz, ntop, omega, lengthsize));
}
// abcei(a, b, c, ei, cc, r1, r2, ibc, nx, ny, nz), main.f:3582
void abcei_(double *a, double *b, double *c, double *ei, const double *cc, const double *r1, const double *r2,
            const int *ibc, const int *nx, const int *ny, const int *nz)
{
    DG_FORM("abcei", dg_abcei(a, b, c, ei, cc, r1, r2, ibc, nx, ny, nz));
}
// bc_scalar(scal, ibc, nx, ny, nz), main.f:4298
void bc_scalar_(double *scal, const int *ibc, const int *nx, const int *ny, const int *nz)
{
    DG_FORM("bc_scalar", dg_bc_scalar(scal, ibc, nx, ny, nz));
}
// filterc27(cc, cfilt, nx, ny, nz), main.f:5473
void filterc27_(const double *cc, double *cfilt, const int *nx, const int *ny, const int *nz)
{
    DG_FORM("filterc27", dg_filterc27(cc, cfilt, nx, ny, nz));
}
// DOUBLE PRECISION FUNCTION AL3D(b1, b2, b3, cc), main.f:4562 (one value per call)
double al3d_(const double *b1, const double *b2, const double *b3, const double *cc)
{
    const int one = 1;
    double out = 0.0;
    DG_FORM("AL3D", dg_al3d_batch(&one, b1, b2, b3, cc, &out));
    return out;
}
// DOUBLE PRECISION FUNCTION FL3D(m1, m2, m3, alpha, r0, dr0), main.f:4639
double fl3d_(const double *m1, const double *m2, const double *m3, const double *alpha, const double *r0,
             const double *dr0)
{
    const int one = 1;
    double out = 0.0;
    DG_FORM("FL3D", dg_fl3d_batch(&one, m1, m2, m3, alpha, r0, dr0, &out));
    return out;
}

#undef DG_FORM

}  // extern "C"
