// dg_prof.h -- optional HIP-event timing of one kernel family on the
// library's own stream, so a caller can measure a kernel's average launch
// duration inside a timed run (dg_profile_start / dg_profile_stop).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace dg {

enum ProfKernel {
    PK_NONE = -1,
    PK_GS_COLOUR = 1,   // tolerance-mode Gauss-Seidel colour pass (all levels, FP32 or FP64)
    PK_RESID = 2,       // tolerance-mode residual (all levels)
    PK_PCG_DIR = 3,     // PCG direction update + matrix-vector product (finest level, FP64)
    PK_PCG_UPDATE = 4,  // PCG x/r update (finest level, FP64)
    PK_RELAX = 5,       // reference relax colour pass (parity mode)
    PK_SWEEP = 6,       // fused two-colour smoother sweep
    PK_RB_COLOUR = 7,   // finest-level colour pass of the red-black split FP32 V-cycle
    PK_CONV_FWD = 8,    // BF16 MFMA convolution forward (the guess network); counts FLOPs, not bytes
};

struct Prof {
    int kernel = PK_NONE;
    std::vector<hipEvent_t> ev;  // pairs
    size_t used = 0;
    double bytes = 0.0;
    long launches = 0;
};

Prof &prof();

inline bool prof_on(int id) { return prof().kernel == id; }

inline void prof_begin(int id, hipStream_t s)
{
    Prof &p = prof();
    if (p.kernel != id) return;
    if (p.used + 2 > p.ev.size()) {
        for (int q = 0; q < 64; q++) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return;
            p.ev.push_back(e);
        }
    }
    (void)hipEventRecord(p.ev[p.used], s);
}

// closes the pair opened by prof_begin; n: the launches of the family the
// pair brackets (a run of consecutive launches shares one pair, so the
// events' own cost is spread over the run)
inline void prof_end(int id, hipStream_t s, double bytes, int n = 1)
{
    Prof &p = prof();
    if (p.kernel != id) return;
    (void)hipEventRecord(p.ev[p.used + 1], s);
    p.used += 2;
    p.bytes += bytes;
    p.launches += n;
}

}  // namespace dg
