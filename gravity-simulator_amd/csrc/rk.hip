// rk.hip -- device-resident Euler, Euler-Cromer, RK4 (src/integrator.c:276-892)
// and the embedded Runge-Kutta methods RKF45, DOPRI, DVERK, RKF78
// (src/integrator_rk_embedded.c:346-987).
//
// The state (x, v, compensated-sum errors, stage derivatives) stays on the
// device; forces come from device_acceleration.  Every per-component update
// is one elementwise FP64 kernel with the reference's operation order
// (contraction off), so with the bitwise FP64 force methods the trajectories
// are the reference's bit for bit.  The embedded methods' error norm is the
// reference's sequential sum over the particles: its terms are formed in
// parallel and added in index order by one lane, then the host runs the
// reference's step-size control on the sum.
#include <cmath>
#include <utility>

#include "gs_internal.h"
#include "rk_tableaus.h"

#pragma clang fp contract(off)

namespace gs {

constexpr int RK_BLOCK = 256;

static dim3 rk_grid(size_t n3) { return dim3(static_cast<unsigned>((n3 + RK_BLOCK - 1) / RK_BLOCK)); }

// src/integrator.c:383-395 (Euler) and :559-570 (Euler-Cromer)
__global__ void k_euler(double *__restrict__ x, double *__restrict__ v,
                        const double *__restrict__ a, double *__restrict__ xe,
                        double *__restrict__ ve, double dt, int cromer, size_t n3)
{
    const size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (q >= n3) return;
    const double x0 = x[q], v0 = v[q];
    double xq, vq, xeq = xe[q], veq = ve[q];
    if (!cromer) {
        xeq += v0 * dt;
        veq += a[q] * dt;
        xq = x0 + xeq;
        vq = v0 + veq;
        xeq += x0 - xq;
        veq += v0 - vq;
    } else {
        veq += a[q] * dt;
        vq = v0 + veq;
        veq += v0 - vq;
        xeq += vq * dt;
        xq = x0 + xeq;
        xeq += x0 - xq;
    }
    x[q] = xq;
    v[q] = vq;
    xe[q] = xeq;
    ve[q] = veq;
}

// RK4 stage state (src/integrator.c:751-797): x = x0 + 0.5 * xk * dt (half
// = 1) or x0 + xk * dt
__global__ void k_rk4_stage(double *__restrict__ x, double *__restrict__ v,
                            const double *__restrict__ x0, const double *__restrict__ v0,
                            const double *__restrict__ xk, const double *__restrict__ vk,
                            double dt, int half, size_t n3)
{
    const size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (q >= n3) return;
    if (half) {
        x[q] = x0[q] + 0.5 * xk[q] * dt;
        v[q] = v0[q] + 0.5 * vk[q] * dt;
    } else {
        x[q] = x0[q] + xk[q] * dt;
        v[q] = v0[q] + vk[q] * dt;
    }
}

// RK4 update with compensated sums (src/integrator.c:799-815); xk, vk hold
// the four stages, n3 apart
__global__ void k_rk4_final(double *__restrict__ x, double *__restrict__ v,
                            const double *__restrict__ x0, const double *__restrict__ v0,
                            const double *__restrict__ xk, const double *__restrict__ vk,
                            double *__restrict__ xe, double *__restrict__ ve, double dt, size_t n3)
{
    const size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (q >= n3) return;
    double veq = ve[q], xeq = xe[q];
    veq += (vk[q] + 2 * vk[n3 + q] + 2 * vk[2 * n3 + q] + vk[3 * n3 + q]) * dt / 6.0;
    xeq += (xk[q] + 2 * xk[n3 + q] + 2 * xk[2 * n3 + q] + xk[3 * n3 + q]) * dt / 6.0;
    const double vq = v0[q] + veq, xq = x0[q] + xeq;
    veq += v0[q] - vq;
    xeq += x0[q] - xq;
    v[q] = vq;
    x[q] = xq;
    ve[q] = veq;
    xe[q] = xeq;
}

struct RkRow {
    double c[RK_MAX_STAGES];
};

// Embedded RK stage s >= 1 (src/integrator_rk_embedded.c:740-781): the
// stage state x + dt * sum_i a[s-1][i] xk_i + xe (likewise v); its velocity
// is the stage's position derivative xk_s
__global__ void k_rke_stage(double *__restrict__ tx, double *__restrict__ xk,
                            const double *__restrict__ vk, const double *__restrict__ x,
                            const double *__restrict__ v, const double *__restrict__ xe,
                            const double *__restrict__ ve, RkRow row, int s, double dt, size_t n3)
{
    const size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (q >= n3) return;
    double sv = 0.0, sx = 0.0;
    for (int i = 0; i < s; i++) {
        sv += row.c[i] * vk[i * n3 + q];
        sx += row.c[i] * xk[i * n3 + q];
    }
    xk[s * n3 + q] = v[q] + dt * sv + ve[q];
    tx[q] = x[q] + dt * sx + xe[q];
}

struct RkWeights {
    double w[RK_MAX_STAGES];
    double dw[RK_MAX_STAGES];  // weights - weights_test
};

// Embedded RK update and error terms (src/integrator_rk_embedded.c:786-853):
// the propagated state with compensated sums into (x1, v1, txe, tve), and
// per component the two squared scaled errors (velocity, then position) in
// the order the reference sums them
__global__ void k_rke_final(double *__restrict__ x1, double *__restrict__ v1,
                            double *__restrict__ txe, double *__restrict__ tve,
                            double *__restrict__ terms, const double *__restrict__ x,
                            const double *__restrict__ v, const double *__restrict__ xe,
                            const double *__restrict__ ve, const double *__restrict__ xk,
                            const double *__restrict__ vk, RkWeights W, int S, double dt,
                            double tol, size_t n3)
{
    const size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (q >= n3) return;
    double sv = 0.0, sx = 0.0, ev = 0.0, ex = 0.0;
    for (int s = 0; s < S; s++) {
        const double vks = vk[s * n3 + q], xks = xk[s * n3 + q];
        sv += W.w[s] * vks;
        sx += W.w[s] * xks;
        ev += dt * W.dw[s] * vks;
        ex += dt * W.dw[s] * xks;
    }
    const double vq = v[q], xq = x[q];
    double tv = ve[q], tx = xe[q];
    tv += dt * sv;
    tx += dt * sx;
    const double v1q = vq + tv, x1q = xq + tx;
    tv += vq - v1q;
    tx += xq - x1q;
    v1[q] = v1q;
    x1[q] = x1q;
    tve[q] = tv;
    txe[q] = tx;
    const double scv = tol + fmax(fabs(vq), fabs(v1q)) * tol;
    const double scx = tol + fmax(fabs(xq), fabs(x1q)) * tol;
    const double e0 = ev / scv, e1 = ex / scx;
    terms[2 * q] = e0 * e0;
    terms[2 * q + 1] = e1 * e1;
}

// Initial-step terms (src/integrator_rk_embedded.c:408-437): per component
// [x^2, v^2] scaled for sum_0 and [v^2, a^2] for sum_1
__global__ void k_rke_init_terms(double *__restrict__ t0, double *__restrict__ t1,
                                 double *__restrict__ tsx, double *__restrict__ tsv,
                                 const double *__restrict__ x, const double *__restrict__ v,
                                 const double *__restrict__ a, double tol, size_t n3)
{
    const size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (q >= n3) return;
    const double sx = tol + tol * fabs(x[q]), sv = tol + tol * fabs(v[q]);
    tsx[q] = sx;
    tsv[q] = sv;
    t0[2 * q] = (x[q] / sx) * (x[q] / sx);
    t0[2 * q + 1] = (v[q] / sv) * (v[q] / sv);
    t1[2 * q] = (v[q] / sx) * (v[q] / sx);
    t1[2 * q + 1] = (a[q] / sv) * (a[q] / sv);
}

// the trial state x + (dt0 / 100) v, v + (dt0 / 100) a (:443-450)
__global__ void k_rke_init_probe(double *__restrict__ x1, double *__restrict__ v1,
                                 const double *__restrict__ x, const double *__restrict__ v,
                                 const double *__restrict__ a, double dt0, size_t n3)
{
    const size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (q >= n3) return;
    x1[q] = x[q] + (dt0 / 100.0) * v[q];
    v1[q] = v[q] + (dt0 / 100.0) * a[q];
}

// sum_2 terms (:466-475)
__global__ void k_rke_init_terms2(double *__restrict__ t2, const double *__restrict__ tsx,
                                  const double *__restrict__ tsv, const double *__restrict__ v,
                                  const double *__restrict__ v1, const double *__restrict__ a,
                                  const double *__restrict__ a1, size_t n3)
{
    const size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (q >= n3) return;
    const double dv = (v1[q] - v[q]) / tsx[q], da = (a1[q] - a[q]) / tsv[q];
    t2[2 * q] = dv * dv;
    t2[2 * q + 1] = da * da;
}

// Sum of `count` terms in index order by one lane (the reference's
// sequential sums, rounding included); the loads run ahead of the adds.
__global__ void k_seq_sum(const double *__restrict__ terms, size_t count, double *__restrict__ out)
{
    if (threadIdx.x != 0) return;
    double s = 0.0;
    size_t k = 0;
    for (; k + 8 <= count; k += 8) {
        double t[8];
#pragma unroll
        for (int j = 0; j < 8; j++) t[j] = terms[k + j];
#pragma unroll
        for (int j = 0; j < 8; j++) s += t[j];
    }
    for (; k < count; k++) s += terms[k];
    *out = s;
}

namespace {

// One device allocation for all the state of a run.
struct RkState {
    Context *c = nullptr;
    double *mem = nullptr;
    size_t n3 = 0;
    double *x = nullptr, *v = nullptr, *m = nullptr;
    ~RkState()
    {
        if (mem) {
            (void)hipStreamSynchronize(c->stream);
            (void)hipFree(mem);
        }
    }
    // `words` doubles beyond x, v (3n each) and m (n), carved by the caller
    ErrorStatus alloc(Context *ctx, int n, size_t words, double **rest)
    {
        c = ctx;
        n3 = static_cast<size_t>(n) * 3;
        const size_t total = 2 * n3 + static_cast<size_t>(n) + words + 2;
        GS_HIP(hipMalloc(reinterpret_cast<void **>(&mem), total * sizeof(double)));
        GS_HIP(hipMemsetAsync(mem, 0, total * sizeof(double), c->stream));
        x = mem;
        v = x + n3;
        m = v + n3;
        *rest = m + n;
        return make_success_error_status();
    }
    ErrorStatus upload(const System *s)
    {
        GS_HIP(hipMemcpyAsync(x, s->x, sizeof(double) * n3, hipMemcpyHostToDevice, c->stream));
        GS_HIP(hipMemcpyAsync(v, s->v, sizeof(double) * n3, hipMemcpyHostToDevice, c->stream));
        GS_HIP(hipMemcpyAsync(m, s->m, sizeof(double) * (n3 / 3), hipMemcpyHostToDevice,
                              c->stream));
        return make_success_error_status();
    }
    ErrorStatus download(System *s, const double *xs, const double *vs)
    {
        GS_HIP(hipMemcpyAsync(s->x, xs, sizeof(double) * n3, hipMemcpyDeviceToHost, c->stream));
        GS_HIP(hipMemcpyAsync(s->v, vs, sizeof(double) * n3, hipMemcpyDeviceToHost, c->stream));
        GS_HIP(hipStreamSynchronize(c->stream));
        return make_success_error_status();
    }
};

ErrorStatus check_args(System *system, AccelerationParam *ap)
{
    if (system->num_particles <= 0)
        return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR, "Number of particles must be positive");
    (void)ap;
    return make_success_error_status();
}

// sequential sum of `count` device terms, read back
ErrorStatus seq_sum(Context *c, const double *terms, size_t count, double *dev_out,
                    double *host_out)
{
    hipLaunchKernelGGL(k_seq_sum, dim3(1), dim3(64), 0, c->stream, terms, count, dev_out);
    GS_HIP(hipGetLastError());
    GS_HIP(hipMemcpyAsync(host_out, dev_out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    return make_success_error_status();
}

}  // namespace

ErrorStatus euler_gpu(System *system, IntegratorParam *ip, AccelerationParam *ap,
                      OutputParam *op, SimulationStatus *status, Settings *settings, double tf,
                      bool cromer)
{
    GS_TRY(check_args(system, ap));
    const int n = system->num_particles;
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    RkState S;
    double *rest;
    const size_t n3 = static_cast<size_t>(n) * 3;
    GS_TRY(S.alloc(c, n, 3 * n3, &rest));
    double *a = rest, *xe = a + n3, *ve = xe + n3;
    GS_TRY(S.upload(system));
    SnapshotWriter snaps;
    GS_TRY(snaps.begin(c, op, system));
    if (snaps.on() && op->output_initial)
        GS_TRY(snaps.write_host(system->x, system->v, status->t, status->dt));
    const double G = system->G;
    double dt = ip->dt;
    const long long total = static_cast<long long>(ceil(tf / dt));
    double t = 0.0;
    long long steps = 0;
    while (steps < total) {
        if (t + dt > tf) dt = tf - t;
        GS_TRY(device_acceleration(c, S.x, S.m, n, G, ap, a));
        hipLaunchKernelGGL(k_euler, rk_grid(n3), dim3(RK_BLOCK), 0, c->stream, S.x, S.v, a, xe,
                           ve, dt, cromer ? 1 : 0, n3);
        GS_HIP(hipGetLastError());
        steps++;
        t = steps * dt;
        if (snaps.due(t)) GS_TRY(snaps.write_device(S.x, S.v, t, dt));
        if (settings->is_exit_ptr && *settings->is_exit_ptr) break;
    }
    GS_TRY(snaps.finish());
    GS_TRY(S.download(system, S.x, S.v));
    status->t = t;
    status->dt = dt;
    status->num_steps = steps;
    return make_success_error_status();
}

ErrorStatus rk4_gpu(System *system, IntegratorParam *ip, AccelerationParam *ap,
                    OutputParam *op, SimulationStatus *status, Settings *settings, double tf)
{
    GS_TRY(check_args(system, ap));
    const int n = system->num_particles;
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    RkState S;
    double *rest;
    const size_t n3 = static_cast<size_t>(n) * 3;
    GS_TRY(S.alloc(c, n, 12 * n3, &rest));
    double *x0 = rest, *v0 = x0 + n3, *xk = v0 + n3, *vk = xk + 4 * n3, *xe = vk + 4 * n3,
           *ve = xe + n3;
    GS_TRY(S.upload(system));
    SnapshotWriter snaps;
    GS_TRY(snaps.begin(c, op, system));
    if (snaps.on() && op->output_initial)
        GS_TRY(snaps.write_host(system->x, system->v, status->t, status->dt));
    const double G = system->G;
    const double dt = ip->dt;
    const long long total = static_cast<long long>(ceil(tf / dt));
    long long steps = 0;
    const size_t b3 = sizeof(double) * n3;
    while (steps < total) {
        GS_HIP(hipMemcpyAsync(x0, S.x, b3, hipMemcpyDeviceToDevice, c->stream));
        GS_HIP(hipMemcpyAsync(v0, S.v, b3, hipMemcpyDeviceToDevice, c->stream));
        for (int k = 0; k < 4; k++) {
            if (k > 0) {
                hipLaunchKernelGGL(k_rk4_stage, rk_grid(n3), dim3(RK_BLOCK), 0, c->stream, S.x,
                                   S.v, x0, v0, xk + (k - 1) * n3, vk + (k - 1) * n3, dt,
                                   k < 3 ? 1 : 0, n3);
                GS_HIP(hipGetLastError());
            }
            GS_TRY(device_acceleration(c, S.x, S.m, n, G, ap, vk + k * n3));
            GS_HIP(hipMemcpyAsync(xk + k * n3, S.v, b3, hipMemcpyDeviceToDevice, c->stream));
        }
        hipLaunchKernelGGL(k_rk4_final, rk_grid(n3), dim3(RK_BLOCK), 0, c->stream, S.x, S.v, x0,
                           v0, xk, vk, xe, ve, dt, n3);
        GS_HIP(hipGetLastError());
        steps++;
        if (snaps.due(steps * dt)) GS_TRY(snaps.write_device(S.x, S.v, steps * dt, dt));
        if (settings->is_exit_ptr && *settings->is_exit_ptr) break;
    }
    GS_TRY(snaps.finish());
    GS_TRY(S.download(system, S.x, S.v));
    status->t = steps * dt;
    status->dt = dt;
    status->num_steps = steps;
    return make_success_error_status();
}

ErrorStatus rk_embedded_gpu(System *system, IntegratorParam *ip, AccelerationParam *ap,
                            OutputParam *op, SimulationStatus *status, Settings *settings,
                            double tf)
{
    GS_TRY(check_args(system, ap));
    // The step control follows the embedded error estimate down to 1e-12 tf:
    // FP32 rounding noise in the forces would drive it there, so the
    // adaptive methods take FP64 forces only.
    if (ap->method == ACCELERATION_METHOD_PAIRWISE_FP32 ||
        ap->method == ACCELERATION_METHOD_MASSLESS_FP32)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR,
                                    "Embedded Runge-Kutta needs FP64 forces; acceleration method "
                                    "%d computes in FP32",
                                    ap->method);
    RkTableau T;
    if (rk_tableau(ip->integrator, &T) != 0)
        return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR, "Invalid integrator method");
    const int n = system->num_particles;
    const int St = T.stages;
    const int min_power = T.power < T.power_test ? T.power : T.power_test;
    RkWeights W;
    for (int s = 0; s < RK_MAX_STAGES; s++) {
        W.w[s] = T.weights[s];
        W.dw[s] = s < St ? T.weights[s] - T.weights_test[s] : 0.0;
    }
    const double tol = ip->tolerance;
    const double safety_max = 6.0, safety_min = 0.33;
    const double safety = pow(0.38, (1.0 / (1.0 + static_cast<double>(min_power))));

    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    RkState S;
    double *rest;
    const size_t n3 = static_cast<size_t>(n) * 3;
    // x1, v1, xe, ve, txe, tve, tx (7 n3), xk, vk (2 St n3), terms (2 n3), sums
    GS_TRY(S.alloc(c, n, (9 + 2 * static_cast<size_t>(St)) * n3 + 8, &rest));
    double *x1 = rest, *v1 = x1 + n3, *xe = v1 + n3, *ve = xe + n3, *txe = ve + n3,
           *tve = txe + n3, *tx = tve + n3, *xk = tx + n3, *vk = xk + St * n3,
           *terms = vk + St * n3, *dsum = terms + 2 * n3;
    void *hp;
    GS_TRY(pinned(c, 64, &hp));
    double *hsum = static_cast<double *>(hp);
    GS_TRY(S.upload(system));
    SnapshotWriter snaps;
    GS_TRY(snaps.begin(c, op, system));
    const double G = system->G;
    double *x = S.x, *v = S.v;
    const dim3 grid = rk_grid(n3), block(RK_BLOCK);

    double dt;
    if (ip->initial_dt > 0.0) {
        dt = ip->initial_dt;
    } else {
        // src/integrator_rk_embedded.c:346-520; the work arrays are free here
        double *a = vk, *a1 = vk + n3, *px = x1, *pv = v1, *tsx = txe, *tsv = tve;
        double *t0 = terms, *t1 = xk;  // (xk holds St >= 6 blocks of n3: room for 2 n3)
        GS_TRY(device_acceleration(c, x, S.m, n, G, ap, a));
        hipLaunchKernelGGL(k_rke_init_terms, grid, block, 0, c->stream, t0, t1, tsx, tsv, x, v, a,
                           tol, n3);
        GS_HIP(hipGetLastError());
        double s0, s1, s2;
        GS_TRY(seq_sum(c, t0, 2 * n3, dsum, hsum));
        s0 = *hsum;
        GS_TRY(seq_sum(c, t1, 2 * n3, dsum, hsum));
        s1 = *hsum;
        const double d0 = sqrt(s0 / (n * 6)), d1 = sqrt(s1 / (n * 6));
        const double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-4 : d0 / d1;
        hipLaunchKernelGGL(k_rke_init_probe, grid, block, 0, c->stream, px, pv, x, v, a, dt0, n3);
        GS_HIP(hipGetLastError());
        GS_TRY(device_acceleration(c, px, S.m, n, G, ap, a1));
        hipLaunchKernelGGL(k_rke_init_terms2, grid, block, 0, c->stream, t0, tsx, tsv, v, pv, a,
                           a1, n3);
        GS_HIP(hipGetLastError());
        GS_TRY(seq_sum(c, t0, 2 * n3, dsum, hsum));
        s2 = *hsum;
        const double d2 = sqrt(s2 / (n * 6)) / dt0;
        // (long-double intermediates rounded to double, as the reference's
        // C arithmetic does before calling pow and fmin)
        const double dt1 = pow(static_cast<double>(0.01L / fmax(d1, d2)),
                               static_cast<double>(1.0L / (1 + T.power)));
        dt = fmin(static_cast<double>(100.0L * dt0), dt1) * 1e-2;
        if (!(dt > 0.0)) return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR, "Initial dt is negative");
        if (dt > tf) dt = tf;
    }

    // initial snapshot after the initial step size (src/integrator_rk_embedded.c:689-718)
    if (snaps.on() && op->output_initial)
        GS_TRY(snaps.write_host(system->x, system->v, status->t, status->dt));
    double t = 0.0;
    long long steps = 0;
    RkRow row;
    while (t < tf) {
        GS_TRY(device_acceleration(c, x, S.m, n, G, ap, vk));
        GS_HIP(hipMemcpyAsync(xk, v, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->stream));
        for (int s = 1; s < St; s++) {
            for (int i = 0; i < RK_MAX_STAGES; i++)
                row.c[i] = i < s ? T.coeff[(s - 1) * (St - 1) + i] : 0.0;
            hipLaunchKernelGGL(k_rke_stage, grid, block, 0, c->stream, tx, xk, vk, x, v, xe, ve,
                               row, s, dt, n3);
            GS_HIP(hipGetLastError());
            GS_TRY(device_acceleration(c, tx, S.m, n, G, ap, vk + s * n3));
        }
        hipLaunchKernelGGL(k_rke_final, grid, block, 0, c->stream, x1, v1, txe, tve, terms, x, v,
                           xe, ve, xk, vk, W, St, dt, tol, n3);
        GS_HIP(hipGetLastError());
        GS_TRY(seq_sum(c, terms, 2 * n3, dsum, hsum));
        double error = sqrt(*hsum / (n * 3 * 2));
        if (error <= 1 || dt <= tf * 1e-12) {
            steps++;
            t += dt;
            std::swap(x, x1);
            std::swap(v, v1);
            std::swap(xe, txe);
            std::swap(ve, tve);
            if (snaps.due(t)) GS_TRY(snaps.write_device(x, v, t, dt));
        }
        if (error < 1e-10) error = 1e-10;
        const double dt_new =
            dt * safety / pow(error, (1.0 / (1.0 + static_cast<double>(min_power))));
        if (dt_new > safety_max * dt)
            dt *= safety_max;
        else if (dt_new < safety_min * dt)
            dt *= safety_min;
        else
            dt = dt_new;
        if (dt_new / tf < 1e-12) dt = tf * 1e-12;
        if (t < tf && t + dt > tf) dt = tf - t;
        if (settings->is_exit_ptr && *settings->is_exit_ptr) break;
    }
    GS_TRY(snaps.finish());
    GS_TRY(S.download(system, x, v));
    status->t = t;
    status->dt = dt;
    status->num_steps = steps;
    return make_success_error_status();
}

}  // namespace gs
