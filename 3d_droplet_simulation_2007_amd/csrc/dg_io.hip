// dg_io.hip -- checkpoint and restart of the resident state.
//
// Two file formats:
//  - the reference's restart file: the records savedata writes
//    (main.f:5414-5467) and iniget reads (main.f:698-793).  They are Fortran
//    unformatted sequential records, each framed by its 4-byte length:
//    t0, nx, ny, nz, tmax, then the inputdata.in values savedata lists, then
//    one 32-byte record per cell (x fastest) holding u, v, w, cc.
//    The reference opens 'saved.dat' on unit 3 but writes unit 2, while
//    iniget reads 'save.dat'.  dg_savedata writes the records to the file
//    it is given, and printout's save at t = tsave writes save.dat.
//    dg_init with restart = 1 reads save.dat beside inputdata.in.  The file
//    holds u, v, w and cc only, so a restart from it is the reference's
//    restart (time, tau and tswap start afresh, stresses at zero), not a
//    continuation of the run.
//  - the full state, for a bit-exact resume (no reference counterpart).  It
//    holds every scalar of the path and every field that carries
//    information from one step to the next: velocity, stresses, densities,
//    viscosities, turbulence fields (with cmu and the Reynolds stresses),
//    interface heights and the previous pressure.  It also holds all eight
//    multigrid heap arrays of every level, because the reference reads
//    stale values of its scratch slots.
//    Layout, little-endian: "DGSTATE1"; int32 nx, ny, nz, ng, nd, ni; nd
//    doubles and ni int32 scalars; an int32 count of arrays; then per array
//    an int32 tag (a DG_F_* id, 999 for cmu, 1000 + 8 * level + slot for the
//    heap) and an int64 length followed by that many doubles.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "droplet_gpu.h"
#include "dg_ctx.h"
#include "dg_resident.h"

namespace dg {

namespace {

constexpr char STATE_MAGIC[8] = {'D', 'G', 'S', 'T', 'A', 'T', 'E', '1'};

// the resident scalars, in file order
void scalars_out(const Scalars &s, std::vector<double> &d, std::vector<int32_t> &i)
{
    d = {s.tau, s.time, s.h, s.tauh, s.rho1, s.rho2, s.mu1, s.mu2, s.sigma, s.g, s.g1, s.xvel1, s.yvel1, s.zvel1,
         s.xvel2, s.yvel2, s.zvel2, s.radius, s.y0, s.z0, s.lengthscale, s.lengthsize, s.timescale, s.vscale,
         s.tau_si, s.maxcfl, s.omega};
    i = {s.tauh2, s.tswap, s.cycles, s.npre, s.ntop, s.npost, s.iturb, s.convectype, s.surfacet, s.hswitch, s.tprint};
}

void scalars_in(Scalars &s, const std::vector<double> &d, const std::vector<int32_t> &i)
{
    size_t q = 0;
    s.tau = d[q++]; s.time = d[q++]; s.h = d[q++]; s.tauh = d[q++]; s.rho1 = d[q++]; s.rho2 = d[q++];
    s.mu1 = d[q++]; s.mu2 = d[q++]; s.sigma = d[q++]; s.g = d[q++]; s.g1 = d[q++]; s.xvel1 = d[q++];
    s.yvel1 = d[q++]; s.zvel1 = d[q++]; s.xvel2 = d[q++]; s.yvel2 = d[q++]; s.zvel2 = d[q++]; s.radius = d[q++];
    s.y0 = d[q++]; s.z0 = d[q++]; s.lengthscale = d[q++]; s.lengthsize = d[q++]; s.timescale = d[q++];
    s.vscale = d[q++]; s.tau_si = d[q++]; s.maxcfl = d[q++]; s.omega = d[q++];
    q = 0;
    s.tauh2 = i[q++]; s.tswap = i[q++]; s.cycles = i[q++]; s.npre = i[q++]; s.ntop = i[q++]; s.npost = i[q++];
    s.iturb = i[q++]; s.convectype = i[q++]; s.surfacet = i[q++]; s.hswitch = i[q++]; s.tprint = i[q++];
}

// the fields a state file holds (absent height and Reynolds-stress arrays are skipped)
const int STATE_FIELDS[] = {DG_F_U, DG_F_V, DG_F_W, DG_F_S11, DG_F_S22, DG_F_S33, DG_F_S12, DG_F_S13, DG_F_S23,
                            DG_F_RHON, DG_F_MUN, DG_F_MUT, DG_F_KTUR, DG_F_EPS, DG_F_PPREV, DG_F_HXMIN, DG_F_HXMAX,
                            DG_F_HYMIN, DG_F_HYMAX, DG_F_HZMIN, DG_F_HZMAX, DG_F_U2P, DG_F_V2P, DG_F_W2P,
                            DG_F_UVP, DG_F_UWP, DG_F_VWP};
constexpr int TAG_CMU = 999;
constexpr int TAG_HEAP = 1000;

bool is_height(int id)
{
    return id == DG_F_HXMIN || id == DG_F_HXMAX || id == DG_F_HYMIN || id == DG_F_HYMAX || id == DG_F_HZMIN ||
           id == DG_F_HZMAX;
}

bool is_reynolds(int id) { return id >= DG_F_U2P && id <= DG_F_VWP; }

// arrays allocated on first use: present in a state file, allocated on load
bool is_lazy(int id) { return is_height(id) || is_reynolds(id); }

// heap slot `slot` of level l, in /pointers/ order: rhs, res, u, cc, a, b, c, ei
double *heap_array(Ctx &c, int l, int slot)
{
    Level &L = c.lev[l];
    double *const a[8] = {L.rhs, L.res, L.u, L.cc, L.a, L.b, L.c, L.ei};
    return a[slot];
}

// device array of a state-file tag (lazily allocated arrays resolve once allocated)
double *tag_array(Ctx &c, int tag, size_t *len)
{
    *len = c.g.n();
    if (tag >= TAG_HEAP) {
        const int l = (tag - TAG_HEAP) / 8, slot = (tag - TAG_HEAP) % 8;
        if (l >= c.ng) return nullptr;
        *len = c.lev[l].g.n();
        return heap_array(c, l, slot);
    }
    if (tag == TAG_CMU) return c.cmu;
    if (tag >= 0 && tag < DG_NFIELDS) return resident_field(tag);
    return nullptr;
}

struct File {
    FILE *f = nullptr;
    ~File()
    {
        if (f) fclose(f);
    }
    // closes the file, reporting a failed flush of buffered data
    int close(const char *path)
    {
        FILE *g = f;
        f = nullptr;
        if (g && fclose(g) != 0) {
            set_error("%s: write failed on close", path);
            return -1;
        }
        return 0;
    }
};

int put(FILE *f, const void *p, size_t n)
{
    if (n && fwrite(p, 1, n, f) != n) {
        set_error("file write failed");
        return -1;
    }
    return 0;
}

int get(FILE *f, void *p, size_t n) { return (n && fread(p, 1, n, f) != n) ? -1 : 0; }

int write_array(FILE *f, int tag, const double *dev, size_t n, std::vector<double> &buf)
{
    buf.resize(n);
    DG_HIP(hipMemcpy(buf.data(), dev, n * sizeof(double), hipMemcpyDeviceToHost));
    const int32_t t = tag;
    const int64_t len = (int64_t)n;
    if (put(f, &t, 4) || put(f, &len, 8) || put(f, buf.data(), n * sizeof(double))) return -1;
    return 0;
}

// ---- Fortran unformatted sequential records --------------------------------
struct RecWriter {
    FILE *f;
    int rec(const void *p, uint32_t n) { return put(f, &n, 4) || put(f, p, n) || put(f, &n, 4) ? -1 : 0; }
    int i(int v) { return rec(&v, 4); }
    int d(double v) { return rec(&v, 8); }
};

struct RecReader {
    FILE *f;
    const char *path;
    int rec(void *p, uint32_t n)
    {
        uint32_t a = 0, b = 0;
        if (get(f, &a, 4)) {
            set_error("restart file %s ends early", path);
            return -1;
        }
        if (a != n) {
            set_error("restart file %s: a record of %u bytes where %u were expected", path, a, n);
            return -1;
        }
        if (get(f, p, n) || get(f, &b, 4)) {
            set_error("restart file %s ends early", path);
            return -1;
        }
        if (b != n) {
            set_error("restart file %s: record trailer %u does not match its length %u", path, b, n);
            return -1;
        }
        return 0;
    }
    int i(int &v) { return rec(&v, 4); }
    int d(double &v) { return rec(&v, 8); }
};

}  // namespace

// iniget (main.f:698-793): the header values iniget reads replace those of
// `in` (tmax becomes t0 + the file's tmax, main.f:791), t0 through *t0, and
// the cell records into uvwcc[0..3] (u, v, w, cc; x fastest).  The grid
// must be that of `in`.
int read_restart_file(const char *path, Input &in, int *t0, std::vector<double> uvwcc[4])
{
    File fh;
    fh.f = fopen(path, "rb");
    if (!fh.f) {
        set_error("restart = %d: cannot open %s (iniget reads save.dat)", in.restart, path);
        return -1;
    }
    RecReader r{fh.f, path};
    int t0v, nx, ny, nz;
    if (r.i(t0v) || r.i(nx) || r.i(ny) || r.i(nz)) return -1;
    if (nx != in.nx || ny != in.ny || nz != in.nz) {
        set_error("restart file %s holds a %d x %d x %d grid, inputdata.in names %d x %d x %d", path, nx, ny, nz,
                  in.nx, in.ny, in.nz);
        return -1;
    }
    Input f = in;
    if (r.i(f.tmax) || r.i(f.nfilter) || r.i(f.cycles) || r.d(f.omega) || r.i(f.npre) || r.i(f.ntop) ||
        r.i(f.npost) || r.i(f.restart) || r.d(f.maxcfl) || r.i(f.iturb) || r.i(f.convectype) || r.i(f.hswitch))
        return -1;
    if (r.d(f.xvel1) || r.d(f.yvel1) || r.d(f.zvel1) || r.d(f.xvel2) || r.d(f.yvel2) || r.d(f.zvel2) ||
        r.d(f.rho2) || r.d(f.rho1) || r.d(f.sigma) || r.d(f.visc2) || r.d(f.visc1) || r.d(f.grav) || r.d(f.grav1))
        return -1;
    if (r.i(f.icase) || r.i(f.ibc) || r.d(f.length) || r.d(f.radius) || r.i(f.tprint) || r.i(f.ntecplot) ||
        r.i(f.tsave))
        return -1;
    if (f.ntecplot == 0) {
        set_error("restart file %s: ntecplot = 0 (iniget STOPs: missing all storing rates, main.f:760-763)", path);
        return -1;
    }
    const size_t n = (size_t)nx * ny * nz;
    for (int q = 0; q < 4; q++) uvwcc[q].resize(n);
    double rec4[4];
    for (size_t p = 0; p < n; p++) {
        if (r.rec(rec4, 32)) return -1;
        for (int q = 0; q < 4; q++) uvwcc[q][p] = rec4[q];
    }
    f.tmax = t0v + f.tmax;
    in = f;
    *t0 = t0v;
    return 0;
}

// savedata's records (main.f:5418-5465) of the run's values and the given
// u, v, w, cc (host arrays, x fastest)
int write_restart_records(const char *path, const Input &in, int t0, const Scalars &s, const double *u,
                          const double *v, const double *w, const double *cc)
{
    File fh;
    fh.f = fopen(path, "wb");
    if (!fh.f) {
        set_error("cannot create %s", path);
        return -1;
    }
    RecWriter r{fh.f};
    if (r.i(t0) || r.i(in.nx) || r.i(in.ny) || r.i(in.nz) || r.i(in.tmax) || r.i(in.nfilter) || r.i(s.cycles) ||
        r.d(s.omega) || r.i(s.npre) || r.i(s.ntop) || r.i(s.npost) || r.i(in.restart) || r.d(s.maxcfl) ||
        r.i(s.iturb) || r.i(s.convectype) || r.i(s.hswitch))
        return -1;
    if (r.d(in.xvel1) || r.d(in.yvel1) || r.d(in.zvel1) || r.d(in.xvel2) || r.d(in.yvel2) || r.d(in.zvel2) ||
        r.d(in.rho2) || r.d(in.rho1) || r.d(in.sigma) || r.d(in.visc2) || r.d(in.visc1) || r.d(in.grav) ||
        r.d(in.grav1))
        return -1;
    if (r.i(in.icase) || r.i(in.ibc) || r.d(in.length) || r.d(in.radius) || r.i(s.tprint) || r.i(in.ntecplot) ||
        r.i(in.tsave))
        return -1;
    const size_t n = (size_t)in.nx * in.ny * in.nz;
    // the cell records, one fwrite per block of cells
    constexpr size_t CELLS = 4096;
    std::vector<unsigned char> blk(CELLS * 40);
    const uint32_t len = 32;
    for (size_t p0 = 0; p0 < n; p0 += CELLS) {
        const size_t p1 = p0 + CELLS < n ? p0 + CELLS : n;
        unsigned char *o = blk.data();
        for (size_t p = p0; p < p1; p++) {
            const double r4[4] = {u[p], v[p], w[p], cc[p]};
            memcpy(o, &len, 4);
            memcpy(o + 4, r4, 32);
            memcpy(o + 36, &len, 4);
            o += 40;
        }
        if (put(fh.f, blk.data(), (size_t)(o - blk.data()))) return -1;
    }
    return fh.close(path);
}

}  // namespace dg

using namespace dg;

extern "C" {

int dg_savedata(const char *path)
{
    if (!resident_ready()) {
        set_error("dg_savedata: dg_init has not been called");
        return -1;
    }
    const RunInfo &ri = run_info();
    if (!ri.valid) {
        set_error("dg_savedata: the run was not set up from an inputdata.in (dg_init), so savedata's values are "
                  "unknown");
        return -1;
    }
    Ctx &c = *ctx();
    DG_HIP(hipStreamSynchronize(c.stream));
    const size_t n = c.g.n();
    std::vector<double> u(n), v(n), w(n), cc(n);
    DG_HIP(hipMemcpy(u.data(), c.u, n * sizeof(double), hipMemcpyDeviceToHost));
    DG_HIP(hipMemcpy(v.data(), c.v, n * sizeof(double), hipMemcpyDeviceToHost));
    DG_HIP(hipMemcpy(w.data(), c.w, n * sizeof(double), hipMemcpyDeviceToHost));
    DG_HIP(hipMemcpy(cc.data(), c.lev[c.ng - 1].cc, n * sizeof(double), hipMemcpyDeviceToHost));
    return write_restart_records(path, ri.in, ri.t0, c.s, u.data(), v.data(), w.data(), cc.data());
}

int dg_run_limits(int *t0, int *tmax)
{
    const RunInfo &ri = run_info();
    if (!resident_ready() || !ri.valid) {
        set_error("dg_run_limits: no run set up from an inputdata.in (dg_init)");
        return -1;
    }
    *t0 = ri.t0;
    *tmax = ri.in.tmax;
    return 0;
}

int dg_save_state(const char *path)
{
    if (!resident_ready()) {
        set_error("dg_save_state: dg_init has not been called");
        return -1;
    }
    Ctx &c = *ctx();
    DG_HIP(hipStreamSynchronize(c.stream));
    File fh;
    fh.f = fopen(path, "wb");
    if (!fh.f) {
        set_error("cannot create %s", path);
        return -1;
    }
    FILE *f = fh.f;
    std::vector<double> d;
    std::vector<int32_t> iv;
    scalars_out(c.s, d, iv);
    const int32_t hdr[6] = {c.g.nx, c.g.ny, c.g.nz, c.ng, (int32_t)d.size(), (int32_t)iv.size()};
    if (put(f, STATE_MAGIC, 8) || put(f, hdr, sizeof hdr) || put(f, d.data(), d.size() * 8) ||
        put(f, iv.data(), iv.size() * 4))
        return -1;
    std::vector<std::pair<int, const double *>> arrs;
    for (int id : STATE_FIELDS) {
        const double *p = resident_field(id);
        if (p) arrs.push_back({id, p});
    }
    arrs.push_back({TAG_CMU, c.cmu});
    const int32_t count = (int32_t)arrs.size() + 8 * c.ng;
    if (put(f, &count, 4)) return -1;
    std::vector<double> buf;
    for (auto &a : arrs)
        if (write_array(f, a.first, a.second, c.g.n(), buf)) return -1;
    for (int l = 0; l < c.ng; l++)
        for (int slot = 0; slot < 8; slot++)
            if (write_array(f, TAG_HEAP + 8 * l + slot, heap_array(c, l, slot), c.lev[l].g.n(), buf)) return -1;
    return fh.close(path);
}

int dg_load_state(const char *path)
{
    if (!resident_ready()) {
        set_error("dg_load_state: dg_init has not been called (a state loads into a run of the same grid)");
        return -1;
    }
    Ctx &c = *ctx();
    DG_HIP(hipStreamSynchronize(c.stream));
    File fh;
    fh.f = fopen(path, "rb");
    if (!fh.f) {
        set_error("cannot open %s", path);
        return -1;
    }
    FILE *f = fh.f;
    char magic[8];
    int32_t hdr[6];
    if (get(f, magic, 8) || memcmp(magic, STATE_MAGIC, 8) != 0) {
        set_error("%s is not a state file of this library", path);
        return -1;
    }
    if (get(f, hdr, sizeof hdr)) {
        set_error("state file %s ends early", path);
        return -1;
    }
    if (hdr[0] != c.g.nx || hdr[1] != c.g.ny || hdr[2] != c.g.nz || hdr[3] != c.ng) {
        set_error("state file %s is for %d x %d x %d (%d levels), the run is %d x %d x %d (%d levels)", path, hdr[0],
                  hdr[1], hdr[2], hdr[3], c.g.nx, c.g.ny, c.g.nz, c.ng);
        return -1;
    }
    std::vector<double> dcur;
    std::vector<int32_t> icur;
    scalars_out(c.s, dcur, icur);
    if ((size_t)hdr[4] != dcur.size() || (size_t)hdr[5] != icur.size()) {
        set_error("state file %s holds %d/%d scalars, this build keeps %zu/%zu", path, hdr[4], hdr[5], dcur.size(),
                  icur.size());
        return -1;
    }
    std::vector<double> d(dcur.size());
    std::vector<int32_t> iv(icur.size());
    int32_t count;
    if (get(f, d.data(), d.size() * 8) || get(f, iv.data(), iv.size() * 4) || get(f, &count, 4)) {
        set_error("state file %s ends early", path);
        return -1;
    }
    // every array is read and checked before anything on the device changes
    std::vector<std::pair<int, std::vector<double>>> arrs;
    for (int q = 0; q < count; q++) {
        int32_t tag;
        int64_t len;
        if (get(f, &tag, 4) || get(f, &len, 8)) {
            set_error("state file %s ends early", path);
            return -1;
        }
        size_t want = 0;
        const bool known = is_lazy(tag) || tag_array(c, tag, &want) != nullptr;
        if (is_lazy(tag)) want = c.g.n();
        if (!known || len < 0 || (size_t)len != want) {
            set_error("state file %s: array tag %d of %lld values does not fit this run", path, tag, (long long)len);
            return -1;
        }
        std::vector<double> data((size_t)len);
        if (get(f, data.data(), (size_t)len * 8)) {
            set_error("state file %s ends early", path);
            return -1;
        }
        arrs.push_back({tag, std::move(data)});
    }
    for (auto &a : arrs)
        if ((is_height(a.first) && alloc_heights(c)) || (is_reynolds(a.first) && alloc_reynolds(c))) return -1;
    for (auto &a : arrs) {
        size_t len = 0;
        double *dst = tag_array(c, a.first, &len);
        DG_HIP(hipMemcpy(dst, a.second.data(), len * sizeof(double), hipMemcpyHostToDevice));
    }
    scalars_in(c.s, d, iv);
    resident_scalars_changed();
    return 0;
}

}  // extern "C"
