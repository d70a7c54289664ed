// context.cpp -- process-wide device context: stream, grow-only buffers,
// pinned staging, kernel timing with HIP events on the library stream.
#include <cstdlib>
#include <cstring>

#include "gs_internal.h"

namespace gs {

namespace {
Context g_ctx;
std::mutex g_init_mu;
bool g_init_done = false;
ErrorStatus g_init_error;
}  // namespace

ErrorStatus context(Context **out)
{
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_init_done) {
        if (g_init_error.return_code != GRAV_SUCCESS)
            return WRAP_RAISE_ERROR(g_init_error.return_code,
                                    "GPU context unavailable (initialisation failed earlier)");
        *out = &g_ctx;
        return make_success_error_status();
    }
    g_init_done = true;
    g_init_error = make_success_error_status();
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
        g_init_error.return_code = GRAV_FAILURE;
        return WRAP_RAISE_ERROR_FMT(GRAV_FAILURE,
                                    "No HIP device available (hipGetDeviceCount: %s, count %d); "
                                    "this library computes on the GPU only",
                                    hipGetErrorString(e), count);
    }
    int dev = 0;
    if (const char *s = getenv("GS_DEVICE")) dev = atoi(s);
    if (dev < 0 || dev >= count) dev = 0;
    e = hipSetDevice(dev);
    if (e == hipSuccess) {
        hipDeviceProp_t prop;
        e = hipGetDeviceProperties(&prop, dev);
        if (e == hipSuccess) {
            g_ctx.cu_count = prop.multiProcessorCount;
            g_ctx.clock_khz = prop.clockRate;
            strncpy(g_ctx.arch, prop.gcnArchName, sizeof(g_ctx.arch) - 1);
            strncpy(g_ctx.name, prop.name, sizeof(g_ctx.name) - 1);
        }
    }
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking);
    // the no-allocation entry point's static scratch: its device address on
    // this device (a failure here is reported by that entry point)
    if (e == hipSuccess) (void)noalloc_scratch_init(dev);
    if (e != hipSuccess) {
        g_init_error.return_code = GRAV_FAILURE;
        return hip_error(e, "device initialisation", __FILE__, __LINE__, __func__);
    }
    g_ctx.device = dev;
    *out = &g_ctx;
    return make_success_error_status();
}

ErrorStatus slot(Context *c, Slot s, size_t bytes, void **out)
{
    Buffer &b = c->slots[s];
    if (bytes == 0) bytes = 256;
    if (b.bytes < bytes) {
        if (b.ptr) {
            GS_HIP(hipStreamSynchronize(c->stream));
            GS_HIP(hipFree(b.ptr));
            b.ptr = nullptr;
            b.bytes = 0;
        }
        size_t want = bytes + bytes / 8;  // headroom against small regrowth
        GS_HIP(hipMalloc(&b.ptr, want));
        b.bytes = want;
    }
    *out = b.ptr;
    return make_success_error_status();
}

ErrorStatus pinned(Context *c, size_t bytes, void **out)
{
    if (c->pinned_bytes < bytes) {
        if (c->pinned) {
            GS_HIP(hipStreamSynchronize(c->stream));
            GS_HIP(hipHostFree(c->pinned));
            c->pinned = nullptr;
            c->pinned_bytes = 0;
        }
        GS_HIP(hipHostMalloc(&c->pinned, bytes, hipHostMallocDefault));
        c->pinned_bytes = bytes;
    }
    *out = c->pinned;
    return make_success_error_status();
}

static int kernel_id(Context *c, const char *name)
{
    for (size_t i = 0; i < c->kernel_names.size(); i++)
        if (c->kernel_names[i] == name) return static_cast<int>(i);
    c->kernel_names.emplace_back(name);
    c->kernel_ms.push_back(0.0);
    c->kernel_calls.push_back(0);
    return static_cast<int>(c->kernel_names.size() - 1);
}

static hipEvent_t take_event(Context *c)
{
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

int timing_begin(Context *c, const char *kernel)
{
    if (!c->timing) return -1;
    Context::Rec r;
    r.kernel = kernel_id(c, kernel);
    r.start = take_event(c);
    r.stop = take_event(c);
    if (!r.start || !r.stop) return -1;
    (void)hipEventRecord(r.start, c->stream);
    c->recs.push_back(r);
    return static_cast<int>(c->recs.size() - 1);
}

void timing_end(Context *c, int rec)
{
    if (rec < 0) return;
    (void)hipEventRecord(c->recs[rec].stop, c->stream);
}

ErrorStatus timing_flush(Context *c)
{
    if (c->recs.empty()) return make_success_error_status();
    GS_HIP(hipStreamSynchronize(c->stream));
    for (auto &r : c->recs) {
        float ms = 0.0f;
        GS_HIP(hipEventElapsedTime(&ms, r.start, r.stop));
        c->kernel_ms[r.kernel] += ms;
        c->kernel_calls[r.kernel] += 1;
        c->event_pool.push_back(r.start);
        c->event_pool.push_back(r.stop);
    }
    c->recs.clear();
    return make_success_error_status();
}

}  // namespace gs

using namespace gs;

extern "C" {

ErrorStatus gs_device_info(GsDeviceInfo *info)
{
    if (!info) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "info is NULL");
    Context *c;
    GS_TRY(context(&c));
    memset(info, 0, sizeof(*info));
    strncpy(info->name, c->name, sizeof(info->name) - 1);
    strncpy(info->arch, c->arch, sizeof(info->arch) - 1);
    info->cu_count = c->cu_count;
    info->clock_khz = c->clock_khz;
    // MI355X published dense vector peaks and HBM3E bandwidth
    info->fp32_vector_tflops = 157.3;
    info->fp64_vector_tflops = 78.6;
    info->hbm_gbs = 8000.0;
    return make_success_error_status();
}

ErrorStatus gs_synchronize(void)
{
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    GS_HIP(hipStreamSynchronize(c->stream));
    return make_success_error_status();
}

ErrorStatus gs_timing_enable(int on)
{
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    c->timing = on != 0;
    return make_success_error_status();
}

ErrorStatus gs_timing_reset(void)
{
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    GS_TRY(timing_flush(c));
    for (auto &v : c->kernel_ms) v = 0.0;
    for (auto &v : c->kernel_calls) v = 0;
    return make_success_error_status();
}

ErrorStatus gs_timing_query(const char *kernel, double *total_ms, long long *calls)
{
    if (!kernel || !total_ms || !calls)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    GS_TRY(timing_flush(c));
    *total_ms = 0.0;
    *calls = 0;
    for (size_t i = 0; i < c->kernel_names.size(); i++)
        if (c->kernel_names[i] == kernel) {
            *total_ms = c->kernel_ms[i];
            *calls = c->kernel_calls[i];
        }
    return make_success_error_status();
}

}  // extern "C"
