// dg_noalloc.hip -- dg_cfl_noalloc: CFL() (main.f:4848-4881) on velocity
// arrays in device memory, allocating nothing.
//
// For callers that cannot allocate device memory: code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc.  The
// reduction's partial maxima live in `scratch`, a file-scope device array
// of a fixed size that every GPU the library runs on holds until the
// process ends.  Host code never uses the array's name as an address: at
// the first call on a device, with that device current, hipGetSymbolAddress
// gives that device's copy, which is kept per device and passed to the
// kernels.  Calls on one GPU share its array: two calls that overlap on
// one GPU race.  Nothing in it needs clearing between calls (every partial
// the final kernel reads is written by the same call).

#include <hip/hip_runtime.h>

#include "droplet_gpu.h"
#include "dg_internal.h"
#include "dg_util.h"

namespace dg {
namespace {

// RED_BLOCKS partial maxima, then the maximum itself
constexpr int SCRATCH_DOUBLES = RED_BLOCKS + 1;
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
double *g_scratch_ptr[MAX_DEVICES] = {nullptr};

// this device's copy of `scratch` (looked up once per device)
int device_scratch(double **out)
{
    int dev = -1;
    DG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) {
        set_error("dg_cfl_noalloc: device %d beyond the %d this build tracks", dev, MAX_DEVICES);
        return -1;
    }
    if (!g_scratch_ptr[dev]) {
        double *p = nullptr;
        DG_HIP(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error("dg_cfl_noalloc: no device address for the scratch array on device %d", dev);
            return -1;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

struct AbsMax3 {
    Grid g;
    const double *u, *v, *w;
    __device__ double operator()(int i, int j, int k) const
    {
        const int o = g.at(i, j, k);
        return max_nan(max_nan(fabs(u[o]), fabs(v[o])), fabs(w[o]));
    }
};

// cfl = max(m, 0) * tauh, as op_cfl forms it on the host
__global__ void k_cfl_scale(const double *m, double tauh, double *cfl)
{
    double v = *m;
    if (v < 0.0) v = 0.0;
    *cfl = v * tauh;
}

}  // namespace
}  // namespace dg

using namespace dg;

extern "C" int dg_cfl_noalloc(const double *u, const double *v, const double *w, const int *nx, const int *ny,
                              const int *nz, const double *tauh, double *cfl, void *stream)
{
    const long long cells = (long long)*nx * *ny * *nz;
    if (*nx < 3 || *ny < 3 || *nz < 3) {
        set_error("dg_cfl_noalloc: extents %d x %d x %d need an interior", *nx, *ny, *nz);
        return -1;
    }
    if (cells > DG_CFL_NOALLOC_MAX_CELLS) {
        set_error("dg_cfl_noalloc: %d x %d x %d = %lld cells exceed the maximum of %lld (32-bit cell indices)", *nx,
                  *ny, *nz, cells, (long long)DG_CFL_NOALLOC_MAX_CELLS);
        return -1;
    }
    if (!u || !v || !w || !cfl) {
        set_error("dg_cfl_noalloc: null array");
        return -1;
    }
    double *sp = nullptr;
    if (device_scratch(&sp)) return -1;
    const Grid g = make_grid(*nx, *ny, *nz);
    hipStream_t s = (hipStream_t)stream;
    // the same two-stage reduction as CFL() on the resident state: at most
    // RED_BLOCKS partials into scratch[0..], their maximum into scratch[RED_BLOCKS]
    if (reduce_interior<true>(g, s, sp, sp + RED_BLOCKS, AbsMax3{g, u, v, w})) return -1;
    hipLaunchKernelGGL(k_cfl_scale, dim3(1), dim3(1), 0, s, sp + RED_BLOCKS, *tauh, cfl);
    DG_CHECK_LAUNCH();
    return 0;
}
