// noalloc.hip -- the direct-sum force evaluation for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).
//
// gs_acceleration_noalloc is acceleration() (methods 1, 2, 16, 17) for
// positions, masses and accelerations that already live in device memory.
// The FP32 direct sum's scratch (per-chunk partial sums and the converted
// source tiles) lives in `scratch`, a file-scope device array of
// NOALLOC_SCRATCH_BYTES that each GPU holds for the life of the process; the
// FP64 sums need none.  Every byte the kernels read from it is written
// earlier in the same call, so nothing has to be cleared.  All calls on one
// GPU share that GPU's array: calls that overlap on one GPU race.
#include <cstdio>
#include <mutex>

#include "gs_internal.h"

namespace gs {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(32) << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;

namespace {
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
}  // namespace

// Device address of `scratch` on the current device (`device`).  In host
// code the array's name only stands in for the symbol, so kernels and copies
// get this address, looked up once per device with that device current.
// Called from the context's initialisation and, for safety, again by the
// entry point (a no-op once the address is known).
hipError_t noalloc_scratch_init(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess) g_scratch_ptr[device] = p;
    return e;
}

static void *noalloc_scratch(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

// The FP32 configuration to use within the scratch: the current one, with
// fewer source chunks when its partial sums would not fit.  Returns false
// when even one chunk does not fit.
static bool fit_fp32_config(int n, int cu_count, Fp32Config &cfg)
{
    if (direct_f32_partial_bytes(n, cfg, cu_count) <= NOALLOC_SCRATCH_BYTES) return true;
    Fp32Config one = cfg;
    one.jsplit = 1;
    const size_t fixed = direct_f32_partial_bytes(n, one, cu_count) - sizeof(double) * 3 * size_t(n);
    if (fixed + sizeof(double) * 3 * size_t(n) > NOALLOC_SCRATCH_BYTES) return false;
    cfg.jsplit = static_cast<int>((NOALLOC_SCRATCH_BYTES - fixed) / (sizeof(double) * 3 * size_t(n)));
    while (cfg.jsplit > 1 && direct_f32_partial_bytes(n, cfg, cu_count) > NOALLOC_SCRATCH_BYTES)
        cfg.jsplit--;
    return direct_f32_partial_bytes(n, cfg, cu_count) <= NOALLOC_SCRATCH_BYTES;
}

}  // namespace gs

using namespace gs;

extern "C" ErrorStatus gs_acceleration_noalloc(double *a, const System *system,
                                               const AccelerationParam *p)
{
    if (!a || !system || !p) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    const int method = p->method;
    if (method != ACCELERATION_METHOD_PAIRWISE && method != ACCELERATION_METHOD_MASSLESS &&
        method != ACCELERATION_METHOD_PAIRWISE_FP32 && method != ACCELERATION_METHOD_MASSLESS_FP32)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR,
                                    "gs_acceleration_noalloc supports the direct sums (methods 1, 2, "
                                    "16, 17) only. Got: %d",
                                    method);
    const int n = system->num_particles;
    if (n <= 0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR,
                                    "Number of particles must be positive. Got: %d", n);
    if (!system->x || !system->m)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "System array x or m is NULL");
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    const double eps = p->softening_length;
    const double G = system->G;
    // the kernels time nothing here: timing would create HIP events
    const bool timing = c->timing;
    c->timing = false;
    hipError_t e = hipSuccess;
    if (method == ACCELERATION_METHOD_PAIRWISE || method == ACCELERATION_METHOD_MASSLESS) {
        e = launch_direct_f64(c, system->x, system->m, n, G, eps,
                              method == ACCELERATION_METHOD_MASSLESS, a);
    } else {
        Fp32Config cfg = fp32_config();
        if (!fit_fp32_config(n, c->cu_count, cfg)) {
            c->timing = timing;
            return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR,
                                        "Number of particles too large for the no-allocation "
                                        "scratch of %zu bytes (FP32 direct sum). Got: %d",
                                        NOALLOC_SCRATCH_BYTES, n);
        }
        e = noalloc_scratch_init(c->device);
        void *buf = noalloc_scratch(c->device);
        if (e == hipSuccess && !buf) e = hipErrorInvalidSymbol;
        if (e == hipSuccess)
            e = launch_direct_f32(c, system->x, system->m, n, G, eps, a, buf,
                                  NOALLOC_SCRATCH_BYTES, cfg);
    }
    c->timing = timing;
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_error(e, "gs_acceleration_noalloc", __FILE__, __LINE__, __func__);
    return make_success_error_status();
}
