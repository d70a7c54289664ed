// output.cpp -- CSV snapshots for the device-resident integrators and the
// reference's output_snapshot entry point (src/output.c:351-469).
//
// The file name (snapshot_%05d.csv in the output directory, numbered by
// OutputParam.output_count_), the four metadata lines, the header and the
// row formats (%d, then %.17g for double and %.8g for float columns) are the
// reference's, character for character, so readers of its CSV files read
// these unchanged.
//
// What differs is where the time goes.  The reference formats every snapshot
// synchronously inside the step loop.  Here the step loop only opens the
// file and queues two device -> pinned-host copies and an event on the run's
// stream; a pool of writer threads waits on the event, formats the rows and
// closes the file while the GPU keeps stepping.  Pinned buffers are recycled
// through a free list: when every buffer is still being formatted, the step
// loop waits for one (the only back-pressure).
#include <sys/stat.h>

#include <algorithm>
#include <charconv>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <string>
#include <thread>
#include <vector>

#include "gs_internal.h"

namespace gs {

namespace {

// Column precisions of src/output.c:425-467: id, m, x, y, z, vx, vy, vz, each
// printed as %.17g (double) or %.8g (float).
struct RowFormat {
    int mass = 17, coord = 17, vel = 17;
};

RowFormat row_format(const OutputParam *op)
{
    RowFormat f;
    if (op->mass_output_dtype == OUTPUT_DTYPE_FLOAT) f.mass = 8;
    if (op->coordinate_output_dtype == OUTPUT_DTYPE_FLOAT) f.coord = 8;
    if (op->velocity_output_dtype == OUTPUT_DTYPE_FLOAT) f.vel = 8;
    return f;
}

std::string snapshot_path(const OutputParam *op)
{
    char name[64];
    snprintf(name, sizeof(name), "snapshot_%05d.csv", op->output_count_);
    return std::string(op->output_dir) + name;
}

// %.<prec>g of d: std::to_chars in the general format with a precision is
// specified to print what printf does, several times faster than snprintf.
inline char *put_g(char *p, char *end, double d, int prec)
{
    return std::to_chars(p, end, d, std::chars_format::general, prec).ptr;
}

// Metadata, header and rows into an open file; false when a write fails.
bool format_snapshot(FILE *f, const RowFormat &fmt, int n, double G, double t, double dt,
                     const int32 *ids, const double *m, const double *x, const double *v)
{
    bool ok = fprintf(f, "# num_particles: %d\n# G: %.17g\n# time: %.17g\n# dt: %.17g\n", n, G, t,
                      dt) > 0;
    ok = ok && fputs("particle_id,m,x,y,z,vx,vy,vz\n", f) >= 0;
    // rows are assembled in a local buffer and written in large blocks; a
    // row is at most 11 + 7 * 24 + 8 characters
    constexpr size_t cap = 1 << 20, row_max = 256;
    std::vector<char> buf(cap);
    char *const b0 = buf.data(), *const end = b0 + cap;
    char *p = b0;
    for (int i = 0; ok && i < n; i++) {
        const double *xi = x + 3 * static_cast<size_t>(i), *vi = v + 3 * static_cast<size_t>(i);
        p = std::to_chars(p, end, ids[i]).ptr;
        *p++ = ',';
        p = put_g(p, end, m[i], fmt.mass);
        for (int d = 0; d < 3; d++) {
            *p++ = ',';
            p = put_g(p, end, xi[d], fmt.coord);
        }
        for (int d = 0; d < 3; d++) {
            *p++ = ',';
            p = put_g(p, end, vi[d], fmt.vel);
        }
        *p++ = '\n';
        if (static_cast<size_t>(end - p) < row_max) {
            ok = fwrite(b0, 1, p - b0, f) == static_cast<size_t>(p - b0);
            p = b0;
        }
    }
    if (ok && p != b0) ok = fwrite(b0, 1, p - b0, f) == static_cast<size_t>(p - b0);
    ok = (fflush(f) == 0) && ok;
    return ok;
}

}  // namespace

ErrorStatus write_csv_snapshot(const char *path, int n, double G, double t, double dt,
                               const int32 *ids, const double *m, const double *x,
                               const double *v, const OutputParam *op)
{
    FILE *f = fopen(path, "w");
    if (!f)
        return WRAP_RAISE_ERROR_FMT(GRAV_OS_ERROR,
                                    "Failed to open file for storing snapshots: \"%s\".", path);
    const bool ok = format_snapshot(f, row_format(op), n, G, t, dt, ids, m, x, v);
    const bool closed = fclose(f) == 0;
    if (!ok || !closed)
        return WRAP_RAISE_ERROR_FMT(GRAV_OS_ERROR, "Failed to write snapshot file: \"%s\".", path);
    return make_success_error_status();
}

// ---------------------------------------------------------------- SnapshotWriter
struct SnapshotWriter::Impl {
    Context *c = nullptr;
    OutputParam *op = nullptr;
    int n = 0;
    size_t n3 = 0;
    double G = 0.0;
    double next = 0.0;  // next output time
    std::vector<int32> ids;
    std::vector<double> m;
    RowFormat fmt;

    struct Buf {
        // pinned: x and v (3 n doubles each), then m (n doubles), ids (n
        // ints) and the count, the last three used by write_device_particles
        double *h = nullptr;
        hipEvent_t ev = nullptr;
    };
    struct Job {
        int buf;
        FILE *f;
        std::string path;
        double t, dt;
        bool device;
        bool particles;  // masses, ids and count are in the buffer
    };
    double *buf_m(int b) const { return bufs[b].h + 2 * n3; }
    int32 *buf_ids(int b) const { return reinterpret_cast<int32 *>(bufs[b].h + 2 * n3 + n); }
    int *buf_count(int b) const { return reinterpret_cast<int *>(buf_ids(b) + n); }
    std::vector<Buf> bufs;
    std::vector<int> free_bufs;
    std::deque<Job> jobs;
    std::mutex mu;
    std::condition_variable cv_job, cv_free;
    std::vector<std::thread> threads;
    bool stop = false;
    std::string error;  // first writer failure
    double *dscratch = nullptr;

    void worker()
    {
        (void)hipSetDevice(c->device);
        for (;;) {
            Job j;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_job.wait(lk, [&] { return stop || !jobs.empty(); });
                if (jobs.empty()) return;
                j = std::move(jobs.front());
                jobs.pop_front();
            }
            std::string err;
            if (j.device) {
                const hipError_t e = hipEventSynchronize(bufs[j.buf].ev);
                if (e != hipSuccess)
                    err = std::string("snapshot copy failed: ") + hipGetErrorString(e);
            }
            if (err.empty()) {
                const double *x = bufs[j.buf].h, *v = x + n3;
                int k = n;
                const int32 *pid = ids.data();
                const double *pm = m.data();
                if (j.particles) {
                    k = *buf_count(j.buf);
                    pid = buf_ids(j.buf);
                    pm = buf_m(j.buf);
                }
                bool ok = k >= 0 && k <= n;
                ok = ok && format_snapshot(j.f, fmt, k, G, j.t, j.dt, pid, pm, x, v);
                if (!ok) err = "failed to write snapshot file \"" + j.path + "\"";
            }
            if (fclose(j.f) != 0 && err.empty())
                err = "failed to close snapshot file \"" + j.path + "\"";
            {
                std::lock_guard<std::mutex> lk(mu);
                if (!err.empty() && error.empty()) error = err;
                free_bufs.push_back(j.buf);
            }
            cv_free.notify_one();
        }
    }

    ErrorStatus pending_error()
    {
        std::lock_guard<std::mutex> lk(mu);
        if (error.empty()) return make_success_error_status();
        return WRAP_RAISE_ERROR_FMT(GRAV_OS_ERROR, "Snapshot output: %s", error.c_str());
    }

    // Opens the next file (caller's thread) and reserves a pinned buffer.
    ErrorStatus open_next(FILE **f, std::string *path, int *buf)
    {
        GS_TRY(pending_error());
        *path = snapshot_path(op);
        *f = fopen(path->c_str(), "w");
        if (!*f)
            return WRAP_RAISE_ERROR_FMT(GRAV_OS_ERROR,
                                        "Failed to open file for storing snapshots: \"%s\".",
                                        path->c_str());
        std::unique_lock<std::mutex> lk(mu);
        cv_free.wait(lk, [&] { return !free_bufs.empty(); });
        *buf = free_bufs.back();
        free_bufs.pop_back();
        return make_success_error_status();
    }

    void submit(int buf, FILE *f, std::string &&path, double t, double dt, bool device,
                bool particles = false)
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            jobs.push_back(Job{buf, f, std::move(path), t, dt, device, particles});
        }
        cv_job.notify_one();
        // src/output.c:331; the step loops then move the next output time
        // (src/integrator.c:415), the initial snapshot does not (:327-340)
        op->output_count_++;
        if (device) next = op->output_count_ * op->output_interval;
    }

    void shutdown()
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv_job.notify_all();
        for (auto &t : threads) t.join();
        threads.clear();
        for (auto &b : bufs) {
            if (b.ev) (void)hipEventDestroy(b.ev);
            if (b.h) (void)hipHostFree(b.h);
        }
        bufs.clear();
        if (dscratch) {
            (void)hipStreamSynchronize(c->stream);
            (void)hipFree(dscratch);
            dscratch = nullptr;
        }
    }
};

SnapshotWriter::~SnapshotWriter()
{
    if (impl_) {
        impl_->shutdown();
        delete impl_;
    }
}

ErrorStatus SnapshotWriter::begin(Context *c, OutputParam *op, const System *sys)
{
    if (!op || op->method == OUTPUT_METHOD_DISABLED) return make_success_error_status();
    if (op->method != OUTPUT_METHOD_CSV)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR,
                                    "Output method %d is not available (CSV only)", op->method);
    if (!op->output_dir)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "Output directory path is NULL.");
    auto *p = new Impl;
    impl_ = p;
    p->c = c;
    p->op = op;
    p->n = sys->num_particles;
    p->n3 = static_cast<size_t>(p->n) * 3;
    p->G = sys->G;
    p->next = op->output_interval;  // src/integrator.c:303
    p->ids.assign(sys->particle_ids, sys->particle_ids + p->n);
    p->m.assign(sys->m, sys->m + p->n);
    p->fmt = row_format(op);
    // Writer threads: formatting runs at several million rows per second per
    // thread; up to 8 threads (half the cores) keep up with snapshots of the
    // largest runs without taking the cores the host side of the run uses.
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nthreads = static_cast<int>(std::min(8u, std::max(1u, hw / 2)));
    // Pinned buffers: one per thread plus two being filled, within 1 GiB.
    const size_t bytes = sizeof(double) * (2 * p->n3 + p->n) + sizeof(int32) * p->n + sizeof(int);
    const size_t by_mem = std::max<size_t>(2, (size_t{1} << 30) / std::max<size_t>(bytes, 1));
    const int nbufs = static_cast<int>(std::min<size_t>(nthreads + 2, by_mem));
    p->bufs.resize(nbufs);
    for (int b = 0; b < nbufs; b++) {
        GS_HIP(hipHostMalloc(reinterpret_cast<void **>(&p->bufs[b].h), bytes,
                             hipHostMallocDefault));
        GS_HIP(hipEventCreateWithFlags(&p->bufs[b].ev, hipEventDisableTiming));
        p->free_bufs.push_back(nbufs - 1 - b);
    }
    for (int t = 0; t < nthreads; t++) p->threads.emplace_back([p] { p->worker(); });
    return make_success_error_status();
}

bool SnapshotWriter::due(double t) const { return impl_ && t >= impl_->next; }

ErrorStatus SnapshotWriter::write_host(const double *x, const double *v, double t, double dt)
{
    if (!impl_) return make_success_error_status();
    Impl *p = impl_;
    FILE *f;
    std::string path;
    int b;
    GS_TRY(p->open_next(&f, &path, &b));
    memcpy(p->bufs[b].h, x, sizeof(double) * p->n3);
    memcpy(p->bufs[b].h + p->n3, v, sizeof(double) * p->n3);
    p->submit(b, f, std::move(path), t, dt, false);
    return make_success_error_status();
}

ErrorStatus SnapshotWriter::write_device(const double *x, const double *v, double t, double dt)
{
    if (!impl_) return make_success_error_status();
    Impl *p = impl_;
    FILE *f;
    std::string path;
    int b;
    GS_TRY(p->open_next(&f, &path, &b));
    const size_t bytes = sizeof(double) * p->n3;
    hipError_t e = hipMemcpyAsync(p->bufs[b].h, x, bytes, hipMemcpyDeviceToHost, p->c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(p->bufs[b].h + p->n3, v, bytes, hipMemcpyDeviceToHost, p->c->stream);
    if (e == hipSuccess) e = hipEventRecord(p->bufs[b].ev, p->c->stream);
    if (e != hipSuccess) {
        fclose(f);
        {
            std::lock_guard<std::mutex> lk(p->mu);
            p->free_bufs.push_back(b);
        }
        GS_HIP(e);
    }
    p->submit(b, f, std::move(path), t, dt, true);
    return make_success_error_status();
}

ErrorStatus SnapshotWriter::write_device_particles(const double *x, const double *v,
                                                   const double *m, const int32 *ids,
                                                   const int *count, double t, double dt)
{
    if (!impl_) return make_success_error_status();
    Impl *p = impl_;
    FILE *f;
    std::string path;
    int b;
    GS_TRY(p->open_next(&f, &path, &b));
    const size_t b3 = sizeof(double) * p->n3;
    hipStream_t st = p->c->stream;
    hipError_t e = hipMemcpyAsync(p->bufs[b].h, x, b3, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(p->bufs[b].h + p->n3, v, b3, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(p->buf_m(b), m, sizeof(double) * p->n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(p->buf_ids(b), ids, sizeof(int32) * p->n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(p->buf_count(b), count, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipEventRecord(p->bufs[b].ev, st);
    if (e != hipSuccess) {
        fclose(f);
        {
            std::lock_guard<std::mutex> lk(p->mu);
            p->free_bufs.push_back(b);
        }
        GS_HIP(e);
    }
    p->submit(b, f, std::move(path), t, dt, true, true);
    return make_success_error_status();
}

ErrorStatus SnapshotWriter::scratch(double **out)
{
    if (!impl_) return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR, "snapshot output is off");
    if (!impl_->dscratch)
        GS_HIP(hipMalloc(reinterpret_cast<void **>(&impl_->dscratch),
                         sizeof(double) * std::max<size_t>(impl_->n3, 1)));
    *out = impl_->dscratch;
    return make_success_error_status();
}

ErrorStatus SnapshotWriter::finish()
{
    if (!impl_) return make_success_error_status();
    impl_->shutdown();
    ErrorStatus st = impl_->pending_error();
    delete impl_;
    impl_ = nullptr;
    return st;
}

}  // namespace gs

extern "C" {

// src/output.c:241-293: one snapshot of the host-side system, synchronously.
ErrorStatus output_snapshot(OutputParam *op, const System *system, const IntegratorParam *ip,
                            const AccelerationParam *ap, const SimulationStatus *status,
                            const Settings *settings)
{
    (void)ip;
    (void)ap;
    (void)settings;
    if (!op || !system || !status)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    switch (op->method) {
        case OUTPUT_METHOD_DISABLED:
            break;
        case OUTPUT_METHOD_CSV: {
            if (!op->output_dir)
                return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "Output directory path is NULL.");
            const std::string path = gs::snapshot_path(op);
            GS_TRY(gs::write_csv_snapshot(path.c_str(), system->num_particles, system->G,
                                          status->t, status->dt, system->particle_ids, system->m,
                                          system->x, system->v, op));
            break;
        }
        case OUTPUT_METHOD_HDF5:
            return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR, "HDF5 output method is not available");
        default:
            return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR, "Unknown output method");
    }
    op->output_count_++;
    return make_success_error_status();
}

}  // extern "C"
