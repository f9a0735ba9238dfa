// dg_output.hip -- printout (main.f:5020-5100) on the resident state: the
// files openfiles creates (main.f:105-118) and printout, velocityc
// (main.f:5773-5808), printtecp (main.f:5106-5153) and savedata
// (main.f:5414-5467) write, in the reference's formats.
//
// A printout does not stop the time step.  On the library's stream, the
// fields it needs are copied into a device snapshot (an HBM-to-HBM copy in
// stream order, so the next step may overwrite the live fields at once).  A
// second stream copies the snapshot to pinned host memory.  A writer thread
// waits for that copy, computes printout's diagnostics from the downloaded
// fields in the reference's loop order (the sums and maxima equal the
// reference's to the bit for equal fields), and writes the files.  Two
// snapshot slots let one printout's download and writing overlap the
// next; a printout that finds both slots busy waits for the older one.
//
// Formats: the .plt records use the reference's explicit edit descriptors
// (1x,e15.7), written as flang's runtime writes Ew.d.  outputdata.out, the
// .plt headers and the screen lines are list-directed; they follow flang's
// list-directed rules: 79-character records, a blank before each item
// except between two character items, REAL(8) as the shortest decimal
// that reads back to the same value, F form when the value rounded to one
// digit lies in [0.1, 1e15], E form otherwise.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "droplet_gpu.h"
#include "dg_ctx.h"
#include "dg_resident.h"

namespace dg {
namespace fmt {

// Non-negative integer in base 1e9 limbs (least significant first), for the
// exact decimal values of doubles.
struct Big {
    std::vector<uint32_t> l;
    static constexpr uint32_t BASE = 1000000000u;
    explicit Big(uint64_t v = 0)
    {
        while (v) {
            l.push_back((uint32_t)(v % BASE));
            v /= BASE;
        }
    }
    void mul(uint32_t m)
    {
        uint64_t c = 0;
        for (uint32_t &d : l) {
            const uint64_t t = (uint64_t)d * m + c;
            d = (uint32_t)(t % BASE);
            c = t / BASE;
        }
        while (c) {
            l.push_back((uint32_t)(c % BASE));
            c /= BASE;
        }
    }
    void mul_pow(uint32_t b, int k)  // *= b^k for b = 2 or 5
    {
        const uint32_t chunk = b == 2 ? (1u << 29) : 1220703125u, cn = b == 2 ? 29 : 13;
        for (; k >= (int)cn; k -= (int)cn) mul(chunk);
        uint32_t r = 1;
        for (; k > 0; k--) r *= b;
        if (r > 1) mul(r);
    }
    void add(const Big &o)
    {
        uint64_t c = 0;
        for (size_t q = 0; q < std::max(l.size(), o.l.size()) || c; q++) {
            if (q == l.size()) l.push_back(0);
            const uint64_t t = (uint64_t)l[q] + (q < o.l.size() ? o.l[q] : 0) + c;
            l[q] = (uint32_t)(t % BASE);
            c = t / BASE;
        }
    }
    bool half()  // /= 2, true when odd
    {
        uint64_t r = 0;
        for (size_t q = l.size(); q-- > 0;) {
            const uint64_t t = r * BASE + l[q];
            l[q] = (uint32_t)(t / 2);
            r = t % 2;
        }
        while (!l.empty() && l.back() == 0) l.pop_back();
        return r != 0;
    }
    void add1()
    {
        Big one(1);
        add(one);
    }
    void sub1()  // > 0
    {
        for (uint32_t &d : l) {
            if (d) {
                d--;
                break;
            }
            d = BASE - 1;
        }
        while (!l.empty() && l.back() == 0) l.pop_back();
    }
    std::string str() const
    {
        if (l.empty()) return "0";
        char b[16];
        snprintf(b, sizeof b, "%u", l.back());
        std::string s = b;
        for (size_t q = l.size() - 1; q-- > 0;) {
            snprintf(b, sizeof b, "%09u", l[q]);
            s += b;
        }
        return s;
    }
};

// a > 0 finite as m * 2^e; the decimal exponent of its exact value as
// m * 5^-e * 10^e (e < 0) or m * 2^e * 10^0
static void split_double(double a, uint64_t *m, int *e)
{
    uint64_t b;
    memcpy(&b, &a, 8);
    const int bexp = (int)((b >> 52) & 0x7ff);
    const uint64_t frac = b & ((1ull << 52) - 1);
    *m = bexp ? frac | (1ull << 52) : frac;
    *e = bexp ? bexp - 1075 : -1074;
}

// the exact value of a (>= 0) as an integer at the decimal scale 10^E
// (E <= a's own decimal exponent)
static Big scaled(double a, int E)
{
    if (a == 0.0) return Big(0);
    uint64_t m;
    int e;
    split_double(a, &m, &e);
    Big v(m);
    int de;
    if (e < 0) {
        v.mul_pow(5, -e);
        de = e;
    } else {
        v.mul_pow(2, e);
        de = 0;
    }
    v.mul_pow(2, de - E);
    v.mul_pow(5, de - E);
    return v;
}

// The decimal digits flang's runtime writes for a > 0 (finite) in
// list-directed output: a = 0.d1d2...dn * 10^e.  The interval is bounded by
// the means with the neighbouring doubles (the lower one rounded up, the
// upper one exclusive); both bounds are split into groups of 16 decimal
// digits aligned at the exponent of the smallest binary digit; the result
// keeps the upper bound's groups down to the first group where the bounds
// differ, drops decimal digits of that group while the bounds still differ
// above the last one (or the last one is 0), and sets the last digit
// halfway between the bounds' digits (rounded up).
static void shortest(double a, char *digits, int *nd, int *e)
{
    const double prev = nextafter(a, 0.0), next = nextafter(a, INFINITY);
    auto dexp = [](double v) {
        uint64_t m;
        int ee;
        split_double(v, &m, &ee);
        return ee < 0 ? ee : 0;
    };
    int E = dexp(a);
    if (prev > 0.0) E = std::min(E, dexp(prev));
    if (std::isfinite(next)) E = std::min(E, dexp(next));
    const Big X = scaled(a, E);
    Big L = scaled(prev, E);
    Big M = std::isfinite(next) ? scaled(next, E) : X;
    L.add(X);
    if (L.half()) L.add1();
    M.add(X);
    if (!M.half()) M.sub1();
    std::string ls = L.str(), ms = M.str();
    // groups of 16 digits counted from the units digit at scale 10^E
    const size_t n = std::max(ls.size(), ms.size());
    const size_t ng = (n + 15) / 16;
    ls.insert(0, ng * 16 - ls.size(), '0');
    ms.insert(0, ng * 16 - ms.size(), '0');
    size_t same = 0;
    while (same < ng && ls.compare(same * 16, 16, ms, same * 16, 16) == 0) same++;
    std::string val;
    int ex;
    if (same == ng) {  // a single representable point: a's own digits
        val = X.str();
        ex = E;
    } else {
        const size_t end = (same + 1) * 16;  // the first differing group ends here
        val = ms.substr(0, end);
        ex = E + (int)((ng - same - 1) * 16);
        uint64_t least = strtoull(ls.substr(same * 16, 16).c_str(), nullptr, 10);
        uint64_t my = strtoull(ms.substr(same * 16, 16).c_str(), nullptr, 10);
        for (;;) {
            const uint64_t q = my / 10, r = my % 10, lq = least / 10, lr = least % 10;
            if (r != 0 && lq == q) {
                val.back() = (char)(val.back() - (char)((r - lr) >> 1));
                break;
            }
            least = lq;
            my = q;
            val.pop_back();
            ex++;
        }
    }
    // leading zeros out, trailing zeros into the exponent
    const size_t first = val.find_first_not_of('0');
    val = first == std::string::npos ? std::string("0") : val.substr(first);
    while (val.size() > 1 && val.back() == '0') {
        val.pop_back();
        ex++;
    }
    if (val.size() > 23) {  // cannot happen for doubles; keeps the buffer bound
        ex += (int)(val.size() - 23);
        val.resize(23);
    }
    memcpy(digits, val.c_str(), val.size() + 1);
    *nd = (int)val.size();
    *e = ex + *nd;
}

// list-directed REAL(8)
std::string list_real(double x)
{
    if (std::isnan(x)) return "NaN";
    if (std::isinf(x)) return x > 0 ? "Inf" : "-Inf";
    std::string s = std::signbit(x) ? "-" : "";
    if (x == 0.0) return s + "0.";
    const double a = fabs(x);
    char one[16];
    snprintf(one, sizeof one, "%.0e", a);
    const int expo1 = atoi(strchr(one, 'e') + 1) + 1;
    char d[24];
    int nd, e;
    shortest(a, d, &nd, &e);
    if (expo1 < 0 || expo1 > 15) {
        s += d[0];
        s += '.';
        s += d + 1;
        char ex[8];
        snprintf(ex, sizeof ex, "E%c%02d", e - 1 < 0 ? '-' : '+', abs(e - 1));
        return s + ex;
    }
    if (e <= 0) return s + "." + std::string((size_t)-e, '0') + d;
    if (e >= nd) return s + d + std::string((size_t)(e - nd), '0') + ".";
    return s + std::string(d, (size_t)e) + "." + (d + e);
}

// E15.7 of a finite nonzero x without printf: the 7 digits are the integer
// nearest to m = |x| 10^(6 - e), formed with one correctly rounded
// multiplication or division by an exact power of ten (relative error below
// 2^-53, so below 1.2e-9 on m).  When m lies within 1e-6 of a rounding tie
// the nearest integer is not certain and the caller takes the exact path;
// likewise for scales beyond 10^22.  Returns false then.
static bool fast_e15_7(double x, char *out)
{
    static const double P10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                                   1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
    const double a = fabs(x);
    int e = (int)floor(log10(a));
    for (int pass = 0; pass < 3; pass++) {
        const int sc = 6 - e;
        if (sc > 22 || sc < -22) return false;
        const double m = sc >= 0 ? a * P10[sc] : a / P10[-sc];
        // the decade with 10^6 <= m < 10^7; a value on the edge of two
        // decades gives the same digits from either (it rounds to 10^6 x
        // 10^k both ways), unless it is a tie, which is tested below
        if (m < 1e6) {
            e -= 1;
            continue;
        }
        if (m >= 1e7) {
            e += 1;
            continue;
        }
        const double fl = floor(m), fr = m - fl;
        if (fabs(fr - 0.5) < 1e-6) return false;
        long long dg = (long long)fl + (fr > 0.5 ? 1 : 0);
        int ex = e + 1;
        if (dg > 9999999) {  // rounded up to 10^7: 0.1000000 of the next decade
            dg /= 10;
            ex += 1;
        }
        char s[24];
        int n = 0;
        if (std::signbit(x)) s[n++] = '-';
        s[n++] = '0';
        s[n++] = '.';
        for (int q = 6; q >= 0; q--) {
            s[n + q] = (char)('0' + dg % 10);
            dg /= 10;
        }
        n += 7;
        const int ae = ex < 0 ? -ex : ex;
        if (ae <= 99) {
            s[n++] = 'E';
            s[n++] = ex < 0 ? '-' : '+';
        } else {
            s[n++] = ex < 0 ? '-' : '+';
            s[n++] = (char)('0' + ae / 100);
        }
        s[n++] = (char)('0' + (ae / 10) % 10);
        s[n++] = (char)('0' + ae % 10);
        memset(out, ' ', (size_t)(15 - n));
        memcpy(out + (15 - n), s, (size_t)n);
        return true;
    }
    return false;
}

// Ew.d, right-justified in w characters (w asterisks when it does not fit).
// exact: always through printf's correctly rounded digits (E15.7 otherwise
// takes fast_e15_7 where it is certain).
void edit_e(char *out, double x, int w, int dd, bool exact = false)
{
    if (!exact && w == 15 && dd == 7 && x != 0.0 && std::isfinite(x) && fast_e15_7(x, out)) return;
    char s[64];
    if (std::isnan(x)) {
        snprintf(s, sizeof s, "NaN");
    } else if (std::isinf(x)) {
        snprintf(s, sizeof s, "%s", x > 0 ? "Inf" : "-Inf");
    } else {
        char digits[40];
        int ex = 0;
        if (x == 0.0) {
            memset(digits, '0', (size_t)dd);
        } else {
            char buf[64];
            snprintf(buf, sizeof buf, "%.*e", dd - 1, fabs(x));
            int n = 0;
            const char *q = buf;
            for (; *q != 'e'; q++)
                if (*q >= '0' && *q <= '9') digits[n++] = *q;
            ex = atoi(q + 1) + 1;
        }
        digits[dd] = 0;
        char eb[8];
        const int ae = abs(ex);
        if (ae <= 99) snprintf(eb, sizeof eb, "E%c%02d", ex < 0 ? '-' : '+', ae);
        else snprintf(eb, sizeof eb, "%c%03d", ex < 0 ? '-' : '+', ae);
        snprintf(s, sizeof s, "%s0.%s%s", std::signbit(x) ? "-" : "", digits, eb);
    }
    const int len = (int)strlen(s);
    if (len > w) {
        memset(out, '*', (size_t)w);
        return;
    }
    memset(out, ' ', (size_t)(w - len));
    memcpy(out + (w - len), s, (size_t)len);
}

// one list-directed WRITE statement
struct ListOut {
    static constexpr size_t RECL = 79;
    std::string s;
    size_t pos = 0;
    bool last_char = false;
    void newline()
    {
        s += '\n';
        pos = 0;
    }
    void put(const char *p, size_t n)
    {
        s.append(p, n);
        pos += n;
    }
    // a numeric item: a blank, then the digits (a new record when they do not fit)
    void num(const std::string &v)
    {
        if (pos > 0 && 1 + v.size() > RECL - pos) newline();
        put(" ", 1);
        put(v.data(), v.size());
        last_char = false;
    }
    // an undelimited character item: a blank unless it follows another
    // character item on the record; split across records when too long
    void chars(const char *c)
    {
        const size_t n = strlen(c);
        const bool space = pos == 0 || !last_char;
        if (pos > 0 && (space ? 1u : 0u) + (n > 0 ? 1u : 0u) > RECL - pos) {
            newline();
        } else if (space) {
            put(" ", 1);
        }
        size_t done = 0;
        while (done < n) {
            const size_t room = RECL - pos;
            if (room == 0) {
                newline();
                put(" ", 1);
                continue;
            }
            const size_t k = std::min(room, n - done);
            put(c + done, k);
            done += k;
        }
        last_char = true;
    }
    ListOut &operator<<(const char *c)
    {
        chars(c);
        return *this;
    }
    ListOut &operator<<(double v)
    {
        num(list_real(v));
        return *this;
    }
    ListOut &operator<<(int v)
    {
        num(std::to_string(v));
        return *this;
    }
    // ends the statement
    std::string end()
    {
        newline();
        return s;
    }
};

// I4.4
std::string i4_4(int v)
{
    if (v < 0 || v > 9999) return "****";
    char b[8];
    snprintf(b, sizeof b, "%04d", v);
    return b;
}

}  // namespace fmt

namespace {

enum { F_U, F_V, F_W, F_CC, F_MUT, F_P, F_N };
enum { W_PRINT = 1, W_TEC = 2, W_PRESS = 4, W_SAVE = 8 };
constexpr int NSLOT = 2;

struct Slot {
    double *dev[F_N] = {};
    double *host[F_N] = {};
    size_t n = 0;
    hipEvent_t snap = nullptr, ready = nullptr;
    bool busy = false;
};

struct Job {
    int slot, t, what;
    unsigned fields;
    Scalars s;
    Input in;
    int t0;
};

struct Writer {
    bool open = false;
    std::string dir;
    bool echo = false;
    FILE *f23 = nullptr, *f24 = nullptr, *f26 = nullptr, *f27 = nullptr, *f28 = nullptr;
    hipStream_t io = nullptr;
    Slot slot[NSLOT];
    int next = 0;
    std::thread th;
    std::mutex m;
    std::condition_variable cv;
    std::deque<Job> q;
    int pending = 0;
    bool stop = false;
    std::string err;  // the writer thread's first error
    // counters
    int jobs = 0;
    double bytes = 0.0, wait_ms = 0.0, write_ms = 0.0;
};

Writer &writer()
{
    static Writer w;
    return w;
}

std::string path_in(const std::string &dir, const std::string &name) { return dir.empty() ? name : dir + "/" + name; }

int fput(FILE *f, const std::string &s)
{
    if (!f) return 0;
    if (fwrite(s.data(), 1, s.size(), f) != s.size()) {
        set_error("output: write failed");
        return -1;
    }
    return 0;
}

// one record of format 111 / 122: n values as (1x,e15.7)
std::string rec_e(const double *v, int n)
{
    std::string s((size_t)(16 * n + 1), ' ');
    for (int q = 0; q < n; q++) fmt::edit_e(&s[(size_t)(16 * q + 1)], v[q], 15, 7);
    s[(size_t)(16 * n)] = '\n';
    return s;
}

// DMAX1 as flang evaluates it: the first argument when it is greater, else the second
inline double dmax1(double a, double b) { return a > b ? a : b; }

// printout's screen and outputdata.out lines, velocityc and the two .plt
// records (main.f:5033-5071)
int write_print(Writer &W, const Job &j, const Slot &sl)
{
    const Input &in = j.in;
    const Scalars &s = j.s;
    const int nx = in.nx, ny = in.ny, nz = in.nz;
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
    const double *u = sl.host[F_U], *v = sl.host[F_V], *w = sl.host[F_W], *cc = sl.host[F_CC],
                 *mut = sl.host[F_MUT];
    auto at = [&](int i, int jj, int k) { return (size_t)(i - 1) + (size_t)(jj - 1) * sy + (size_t)(k - 1) * sz; };
    // velocityc (main.f:5773-5808): centroid of cc on the plane k = (nz-2)/2
    {
        const int k = (nz - 2) / 2;
        double sum = 0.0, ii = 0.0, jj = 0.0;
        for (int i = 2; i <= nx - 1; i++)
            for (int q = 2; q <= ny - 1; q++) {
                ii = (double)i * cc[at(i, q, k)] + ii;
                jj = (double)q * cc[at(i, q, k)] + jj;
            }
        for (int i = 2; i <= nx - 1; i++)
            for (int q = 2; q <= ny - 1; q++) sum = cc[at(i, q, k)] + sum;
        ii = ii / sum;
        jj = jj / sum;
        double vvv = NAN;
        // the reference indexes u(i+1, j+1) with i = int(ii), j = int(jj);
        // a centroid outside the grid (no liquid on the plane) gives NaN here
        if (std::isfinite(ii) && std::isfinite(jj) && ii >= 1.0 && jj >= 1.0 && ii < nx && jj < ny && k >= 1) {
            const int i = (int)ii, q = (int)jj;
            const double fi = ii - (double)i, fj = jj - (double)q;
            const double uuu = u[at(i, q + 1, k)] + (u[at(i + 1, q + 1, k)] - u[at(i, q + 1, k)]) * fi;
            const double uud = u[at(i, q, k)] + (u[at(i + 1, q, k)] - u[at(i, q, k)]) * fi;
            const double uu = uud + (uuu - uud) * fj;
            const double vvu = v[at(i, q + 1, k)] + (v[at(i + 1, q + 1, k)] - v[at(i, q + 1, k)]) * fi;
            const double vvd = v[at(i, q, k)] + (v[at(i + 1, q, k)] - v[at(i, q, k)]) * fi;
            const double vv = vvd + (vvu - vvd) * fj;
            vvv = sqrt(vv * vv + uu * uu);
        }
        const double r[2] = {s.time * s.timescale * 1000.0, vvv * s.lengthscale / s.timescale};
        if (fput(W.f28, rec_e(r, 2))) return -1;
    }
    // CFL() (main.f:4848-4881), FINDNINF, SUMFIELD, FINDMAX, FINDN2 (main.f:4700-4846)
    double cfl = 0.0;
    for (const double *f : {u, v, w})
        for (int k = 2; k <= nz - 1; k++)
            for (int q = 2; q <= ny - 1; q++)
                for (int i = 2; i <= nx - 1; i++) cfl = dmax1(cfl, fabs(f[at(i, q, k)]));
    cfl = cfl * s.tauh;
    double velmax = 0.0, totalcc = 0.0, totalxmom = 0.0, maxmut = mut[at(2, 2, 2)], meanmut = 0.0, n2 = 0.0;
    for (int k = 2; k <= nz - 1; k++)
        for (int q = 2; q <= ny - 1; q++)
            for (int i = 2; i <= nx - 1; i++) {
                const size_t o = at(i, q, k);
                const double vv = sqrt(u[o] * u[o] + v[o] * v[o] + w[o] * w[o]);
                velmax = dmax1(vv, velmax);
                totalcc = totalcc + cc[o];
                totalxmom = totalxmom + u[o];
                if (mut[o] > maxmut) maxmut = mut[o];
                meanmut = meanmut + mut[o];
                n2 = n2 + u[o] * u[o] + v[o] * v[o] + w[o] * w[o];
            }
    maxmut = maxmut / s.mu1;
    meanmut = meanmut / (double)((nx - 2) * (ny - 2) * (nz - 2)) / s.mu1;
    const double tnpt = (double)((nx - 2) * (ny - 2) * (nz - 2));
    const double norm = sqrt(n2 / tnpt);
    const double tsi = s.time * s.timescale;

    // the star lines carry the blank that column 72 of the source holds
    const std::string stars6 = std::string(51, '*') + " " + std::string(25, '*');
    const std::string stars23 = std::string(50, '*') + " " + std::string(25, '*');
    std::string scr, o23;
    scr += (fmt::ListOut() << stars6.c_str()).end();
    o23 += (fmt::ListOut() << stars23.c_str()).end();
    scr += (fmt::ListOut() << " timestep=" << j.t).end();
    scr += (fmt::ListOut() << " time=" << s.time << "  time_SI=" << tsi).end();
    o23 += (fmt::ListOut() << " timestep=" << j.t).end();
    o23 += (fmt::ListOut() << " time=" << s.time << "  time_SI=" << tsi).end();
    scr += (fmt::ListOut() << " tau=" << s.tau << "   tau_SI=" << s.tau * s.timescale).end();
    o23 += (fmt::ListOut() << "  tau=" << s.tau << "   tau_SI=" << s.tau * s.timescale).end();
    scr += (fmt::ListOut() << " CFL = " << cfl).end();
    o23 += (fmt::ListOut() << " CFL = " << cfl).end();
    scr += (fmt::ListOut() << "max |v|= " << velmax << "total cc = " << totalcc).end();
    o23 += (fmt::ListOut() << "max |v|= " << velmax << "total cc = " << totalcc).end();
    scr += (fmt::ListOut() << "total x momentum= " << totalxmom).end();
    o23 += (fmt::ListOut() << "total x momentum= " << totalxmom).end();
    scr += (fmt::ListOut() << "max mut/mu1=" << maxmut << " mean mut/mu1=" << meanmut).end();
    o23 += (fmt::ListOut() << "max mut/mu1=" << maxmut << " mean mut/mu1=" << meanmut).end();
    if (W.echo) fput(stdout, scr);
    if (fput(W.f23, o23)) return -1;
    const double r24[2] = {tsi, velmax * s.lengthscale / s.timescale};
    const double r26[2] = {tsi, norm * s.lengthscale / s.timescale};
    if (fput(W.f24, rec_e(r24, 2)) || fput(W.f26, rec_e(r26, 2))) return -1;
    return 0;
}

// printtecp (main.f:5106-5153): mainNNNN.plt, NNNN = t/ntecplot; records
// formatted by several threads, written in order
int write_tecplot(Writer &W, const Job &j, const Slot &sl)
{
    const Input &in = j.in;
    const Scalars &s = j.s;
    const int nx = in.nx, ny = in.ny, nz = in.nz;
    if (W.echo) fput(stdout, (fmt::ListOut() << "ntecplot print, t=" << j.t).end());
    const std::string name = path_in(W.dir, "main" + fmt::i4_4(j.t / in.ntecplot) + ".plt");
    FILE *f = fopen(name.c_str(), "wb");
    if (!f) {
        set_error("output: cannot create %s", name.c_str());
        return -1;
    }
    std::string head = (fmt::ListOut() << "VARIABLES=\"X\",\"Y\",\"Z\",\"U\",\"V\",\"W\",\"CC\"").end();
    head += (fmt::ListOut() << "zone i=" << nx << " j=" << ny - 1 << " k=" << nz).end();
    int rc = fput(f, head);
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
    const double *u = sl.host[F_U], *v = sl.host[F_V], *w = sl.host[F_W], *cc = sl.host[F_CC];
    constexpr size_t LINE = 7 * 16 + 1;
    auto plane = [&](int k, std::string &buf) {
        buf.resize((size_t)(ny - 1) * nx * LINE);
        char *o = &buf[0];
        const double zz = (double)k * s.h * s.lengthscale * 1000.0;
        for (int jj = 2; jj <= ny; jj++) {
            const double yy = (double)jj * s.h * s.lengthscale * 1000.0;
            for (int i = 1; i <= nx; i++) {
                const double xx = (double)i * s.h * s.lengthscale * 1000.0;
                const size_t c0 = (size_t)(i - 1) + (size_t)(jj - 1) * sy + (size_t)(k - 1) * sz;
                const double un = i == nx ? u[c0] : 0.5 * (u[c0] + u[c0 + 1]);
                const double vn = jj == ny ? v[c0] : 0.5 * (v[c0] + v[c0 + sy]);
                const double wn = k == nz ? w[c0] : 0.5 * (w[c0] + w[c0 + sz]);
                const double r[7] = {xx, yy, zz, un, vn, wn, cc[c0]};
                *o = ' ';
                for (int q = 0; q < 7; q++) {
                    o[16 * q] = ' ';
                    fmt::edit_e(o + 16 * q + 1, r[q], 15, 7);
                }
                o[7 * 16] = '\n';
                o += LINE;
            }
        }
    };
    // formatting threads: the host's cores but two, kept for the thread
    // that queues the time steps and for the runtime's own
    const unsigned hw = std::thread::hardware_concurrency();
    const unsigned nth = std::max(1u, std::min(hw > 2 ? hw - 2 : 1u, 16u));
    // planes in batches of nth, each plane formatted by one thread
    std::vector<std::string> bufs(nth);
    for (int k0 = 1; k0 <= nz && rc == 0; k0 += (int)nth) {
        const int kn = std::min((int)nth, nz - k0 + 1);
        std::vector<std::thread> ths;
        for (int q = 1; q < kn; q++) ths.emplace_back(plane, k0 + q, std::ref(bufs[(size_t)q]));
        plane(k0, bufs[0]);
        for (auto &t : ths) t.join();
        for (int q = 0; q < kn && rc == 0; q++) rc = fput(f, bufs[(size_t)q]);
    }
    if (fclose(f) != 0 && rc == 0) {
        set_error("output: write failed on %s", name.c_str());
        rc = -1;
    }
    return rc;
}

// pressureNNNN.plt (main.f:5078-5095), NNNN = t/100: the plane k = (nz-2)/2
int write_pressure(Writer &W, const Job &j, const Slot &sl)
{
    const Input &in = j.in;
    const Scalars &s = j.s;
    const int nx = in.nx, ny = in.ny, nz = in.nz;
    const std::string name = path_in(W.dir, "pressure" + fmt::i4_4(j.t / 100) + ".plt");
    FILE *f = fopen(name.c_str(), "wb");
    if (!f) {
        set_error("output: cannot create %s", name.c_str());
        return -1;
    }
    std::string out = (fmt::ListOut() << "VARIABLES=\"x\",\"y\",\"p\"").end();
    out += (fmt::ListOut() << "zone i=" << nx - 2 << "j=" << ny - 2).end();
    const int k = (nz - 2) / 2;
    const double *p = sl.host[F_P];
    for (int jj = 2; jj <= ny - 1; jj++)
        for (int i = 2; i <= nx - 1; i++) {
            const double yy = s.h * (double)jj, xx = s.h * (double)i;
            const double pn = k >= 1 ? p[(size_t)(i - 1) + (size_t)(jj - 1) * nx + (size_t)(k - 1) * nx * ny] /
                                           (s.tauh * s.tauh) * in.rho1
                                     : NAN;
            const double r[3] = {xx, yy, pn};
            out += rec_e(r, 3);
        }
    int rc = fput(f, out);
    if (fclose(f) != 0 && rc == 0) {
        set_error("output: write failed on %s", name.c_str());
        rc = -1;
    }
    return rc;
}

void writer_loop(int device)
{
    Writer &W = writer();
    std::string err;
    set_error_sink(&err);
    hipSetDevice(device);
    for (;;) {
        Job j;
        {
            std::unique_lock<std::mutex> lk(W.m);
            W.cv.wait(lk, [&] { return W.stop || !W.q.empty(); });
            if (W.q.empty()) break;
            j = W.q.front();
            W.q.pop_front();
        }
        Slot &sl = W.slot[j.slot];
        const auto t0 = std::chrono::steady_clock::now();
        err.clear();
        if (hipEventSynchronize(sl.ready) != hipSuccess) set_error("output: download of step %d failed", j.t);
        if (err.empty() && (j.what & W_PRINT)) write_print(W, j, sl);
        if (err.empty() && (j.what & W_TEC)) write_tecplot(W, j, sl);
        if (err.empty() && (j.what & W_PRESS)) write_pressure(W, j, sl);
        if (err.empty() && (j.what & W_SAVE)) {
            const std::string name = path_in(W.dir, "save.dat");
            write_restart_records(name.c_str(), j.in, j.t0, j.s, sl.host[F_U], sl.host[F_V], sl.host[F_W],
                                  sl.host[F_CC]);
        }
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        std::lock_guard<std::mutex> lk(W.m);
        if (!err.empty() && W.err.empty()) W.err = "printout of step " + std::to_string(j.t) + ": " + err;
        W.write_ms += ms;
        sl.busy = false;
        W.pending--;
        W.cv.notify_all();
    }
    set_error_sink(nullptr);
}

// waits until the writer has finished every queued printout
int drain(Writer &W)
{
    std::unique_lock<std::mutex> lk(W.m);
    W.cv.wait(lk, [&] { return W.pending == 0; });
    if (!W.err.empty()) {
        set_error("%s", W.err.c_str());
        W.err.clear();
        return -1;
    }
    return 0;
}

void close_files(Writer &W)
{
    for (FILE **f : {&W.f23, &W.f24, &W.f26, &W.f27, &W.f28})
        if (*f) {
            fclose(*f);
            *f = nullptr;
        }
    W.open = false;
}

void free_slots(Writer &W)
{
    for (Slot &sl : W.slot) {
        for (int q = 0; q < F_N; q++) {
            if (sl.dev[q]) hipFree(sl.dev[q]);
            if (sl.host[q]) hipHostFree(sl.host[q]);
            sl.dev[q] = sl.host[q] = nullptr;
        }
        if (sl.snap) hipEventDestroy(sl.snap);
        if (sl.ready) hipEventDestroy(sl.ready);
        sl.snap = sl.ready = nullptr;
        sl.n = 0;
        sl.busy = false;
    }
    if (W.io) hipStreamDestroy(W.io);
    W.io = nullptr;
}

}  // namespace

bool output_active() { return writer().open; }

int output_shutdown()
{
    Writer &W = writer();
    int rc = 0;
    if (W.th.joinable()) {
        rc = drain(W);
        {
            std::lock_guard<std::mutex> lk(W.m);
            W.stop = true;
        }
        W.cv.notify_all();
        W.th.join();
        W.stop = false;
    }
    close_files(W);
    free_slots(W);
    return rc;
}

}  // namespace dg

using namespace dg;

extern "C" {

int dg_output_open(const char *dir, const int *echo_stdout)
{
    Writer &W = writer();
    if (W.th.joinable() && drain(W)) return -1;
    close_files(W);
    W.dir = dir ? dir : "";
    W.echo = echo_stdout && *echo_stdout != 0;
    // openfiles (main.f:105-118)
    struct {
        FILE **f;
        const char *name, *head;
    } files[] = {{&W.f23, "outputdata.out", nullptr},
                 {&W.f24, "spurious.plt", "VARIABLES=\"time\",\"Vmax\""},
                 {&W.f26, "normv.plt", "VARIABLES=\"time\",\"norm V\""},
                 {&W.f27, "hysteresis.plt", "VARIABLES=\"time\",\"thetaH\",\"theta A\",\"theta R\""},
                 {&W.f28, "velocity.plt", "VARIABLES=\"time (ms)\",\"u-center\""}};
    for (auto &e : files) {
        const std::string name = path_in(W.dir, e.name);
        *e.f = fopen(name.c_str(), "wb");
        if (!*e.f) {
            set_error("dg_output_open: cannot create %s", name.c_str());
            close_files(W);
            return -1;
        }
    }
    for (auto &e : files)
        if (e.head && fput(*e.f, (fmt::ListOut() << e.head).end())) {
            close_files(W);
            return -1;
        }
    W.open = true;
    if (!W.th.joinable()) {
        int dev = 0;
        DG_HIP(hipGetDevice(&dev));
        W.th = std::thread(writer_loop, dev);
    }
    return 0;
}

int dg_printout(const int *t)
{
    if (!resident_ready()) {
        set_error("dg_printout: dg_init has not been called");
        return -1;
    }
    const RunInfo &ri = run_info();
    if (!ri.valid) {
        set_error("dg_printout: the run was not set up from an inputdata.in (dg_init)");
        return -1;
    }
    Writer &W = writer();
    if (!W.open) {
        set_error("dg_printout: dg_output_open has not been called");
        return -1;
    }
    Ctx &c = *ctx();
    const Input &in = ri.in;
    const int tt = *t;
    int what = 0;
    if (c.s.tprint > 0 && tt % c.s.tprint == 0) what |= W_PRINT;
    if (in.ntecplot > 0 && tt % in.ntecplot == 0) what |= W_TEC;
    if (tt % 100 == 0) what |= W_PRESS;
    if (tt == in.tsave) what |= W_SAVE;
    if (!what) return 0;
    unsigned fields = 0;
    if (what & W_PRINT) fields |= 1u << F_U | 1u << F_V | 1u << F_W | 1u << F_CC | 1u << F_MUT;
    if (what & (W_TEC | W_SAVE)) fields |= 1u << F_U | 1u << F_V | 1u << F_W | 1u << F_CC;
    if (what & W_PRESS) fields |= 1u << F_P;

    // a free slot (the older printout's download and files finish first)
    Slot &sl = W.slot[W.next];
    {
        const auto w0 = std::chrono::steady_clock::now();
        std::unique_lock<std::mutex> lk(W.m);
        W.cv.wait(lk, [&] { return !sl.busy; });
        W.wait_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
        if (!W.err.empty()) {
            set_error("%s", W.err.c_str());
            W.err.clear();
            return -1;
        }
    }
    const size_t n = c.g.n(), bytes = n * sizeof(double);
    if (sl.n != n) {
        for (int q = 0; q < F_N; q++) {
            if (sl.dev[q]) DG_HIP(hipFree(sl.dev[q]));
            if (sl.host[q]) DG_HIP(hipHostFree(sl.host[q]));
            sl.dev[q] = sl.host[q] = nullptr;
        }
        sl.n = n;
    }
    if (!W.io) DG_HIP(hipStreamCreateWithFlags(&W.io, hipStreamNonBlocking));
    if (!sl.snap) DG_HIP(hipEventCreateWithFlags(&sl.snap, hipEventDisableTiming));
    if (!sl.ready) DG_HIP(hipEventCreateWithFlags(&sl.ready, hipEventDisableTiming));
    const double *src[F_N] = {c.u, c.v, c.w, c.lev[c.ng - 1].cc, c.mut, c.lev[c.ng - 1].u};
    double moved = 0.0;
    for (int q = 0; q < F_N; q++) {
        if (!(fields & (1u << q))) continue;
        if (!sl.dev[q]) DG_HIP(hipMalloc(&sl.dev[q], bytes));
        if (!sl.host[q]) DG_HIP(hipHostMalloc(&sl.host[q], bytes, hipHostMallocDefault));
        DG_HIP(hipMemcpyAsync(sl.dev[q], src[q], bytes, hipMemcpyDeviceToDevice, c.stream));
    }
    DG_HIP(hipEventRecord(sl.snap, c.stream));
    DG_HIP(hipStreamWaitEvent(W.io, sl.snap, 0));
    for (int q = 0; q < F_N; q++) {
        if (!(fields & (1u << q))) continue;
        DG_HIP(hipMemcpyAsync(sl.host[q], sl.dev[q], bytes, hipMemcpyDeviceToHost, W.io));
        moved += (double)bytes;
    }
    DG_HIP(hipEventRecord(sl.ready, W.io));
    {
        std::lock_guard<std::mutex> lk(W.m);
        sl.busy = true;
        W.pending++;
        W.jobs++;
        W.bytes += moved;
        W.q.push_back(Job{W.next, tt, what, fields, c.s, in, ri.t0});
    }
    W.cv.notify_all();
    W.next = (W.next + 1) % NSLOT;
    return 0;
}

int dg_output_flush(void)
{
    Writer &W = writer();
    if (W.th.joinable() && drain(W)) return -1;
    for (FILE *f : {W.f23, W.f24, W.f26, W.f27, W.f28})
        if (f && fflush(f) != 0) {
            set_error("dg_output_flush: write failed");
            return -1;
        }
    fflush(stdout);
    return 0;
}

int dg_output_close(void) { return output_shutdown(); }

int dg_output_stats(int *jobs, double *bytes, double *wait_ms, double *write_ms)
{
    Writer &W = writer();
    std::lock_guard<std::mutex> lk(W.m);
    if (jobs) *jobs = W.jobs;
    if (bytes) *bytes = W.bytes;
    if (wait_ms) *wait_ms = W.wait_ms;
    if (write_ms) *write_ms = W.write_ms;
    return 0;
}

int dg_format_real(const double *x, const int *mode, char *buf, const int *len)
{
    std::string s;
    if (*mode == 0) {
        s = fmt::list_real(*x);
    } else if (*mode == 1 || *mode == 2) {
        s.assign(15, ' ');
        fmt::edit_e(&s[0], *x, 15, 7, *mode == 2);
    } else {
        set_error("dg_format_real: mode %d (0 list-directed, 1 E15.7, 2 E15.7 through printf's digits)", *mode);
        return -1;
    }
    if ((int)s.size() + 1 > *len) {
        set_error("dg_format_real: buffer of %d bytes is too small", *len);
        return -1;
    }
    memcpy(buf, s.c_str(), s.size() + 1);
    return 0;
}

}  // extern "C"
