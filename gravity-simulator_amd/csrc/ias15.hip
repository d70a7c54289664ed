// ias15.hip -- IAS15, the 15th-order Gauss-Radau predictor-corrector with
// adaptive steps (src/integrator_ias15.c:229-1382), device-resident on the
// GPU.  Every per-component update (predictor, g/b recurrences, step end,
// b refinement) is one elementwise FP64 kernel with the reference's operation
// order; the force evaluations go through device_acceleration, so with the
// bitwise FP64 pairwise sum the run reproduces the reference bit for bit.
// The host keeps the reference's scalar control flow (convergence test,
// step-size control) and reads back two block-wise maxima per test.
#include <algorithm>
#include <cmath>
#include <vector>

#include "gs_internal.h"
#include "ias15_coeffs.h"

#pragma clang fp contract(off)

namespace gs {

constexpr int IAS_BLOCK = 256;
constexpr int IAS_MAX_GRID = 1024;

struct IasK {
    double c[7][7];
    double r[8][8];
};

#define AT(arr, slot) (arr)[static_cast<size_t>(slot) * n3 + q]

// predictor at spacing h (src/integrator_ias15.c:808-919)
__global__ void k_ias_predict(double *__restrict__ x1, double *__restrict__ v1,
                              const double *__restrict__ x, const double *__restrict__ v,
                              const double *__restrict__ a, const double *__restrict__ b,
                              const double *__restrict__ xc, const double *__restrict__ vc,
                              double h, double dt, size_t n3)
{
    const size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (q >= n3) return;
    const double b0 = AT(b, 0), b1 = AT(b, 1), b2 = AT(b, 2), b3 = AT(b, 3), b4 = AT(b, 4),
                 b5 = AT(b, 5), b6 = AT(b, 6);
    double px = b5 / 28.0 + h * b6 / 36.0;
    px = b4 / 21.0 + h * px;
    px = b3 / 15.0 + h * px;
    px = b2 / 10.0 + h * px;
    px = b1 / 6.0 + h * px;
    px = b0 / 3.0 + h * px;
    px = a[q] + h * px;
    double xn = x[q];
    xn += xc[q] + dt * h * (v[q] + dt * h * px / 2.0);
    x1[q] = xn;

    double pv = b5 / 7.0 + h * b6 / 8.0;
    pv = b4 / 6.0 + h * pv;
    pv = b3 / 5.0 + h * pv;
    pv = b2 / 4.0 + h * pv;
    pv = b1 / 3.0 + h * pv;
    pv = b0 / 2.0 + h * pv;
    double vn = v[q];
    vn += vc[q] + dt * h * (a[q] + h * pv);
    v1[q] = vn;
}

// g (divided differences of the forces F at spacings 0..i) and b after the
// force at spacing i (src/integrator_ias15.c:1117-1274, :985-1115)
__global__ void k_ias_gb(double *__restrict__ g, double *__restrict__ b,
                         const double *__restrict__ F, int i, IasK k, size_t n3)
{
    const size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (q >= n3) return;
    const int m = i < 7 ? i : 7;
    double gl[7];
#pragma unroll
    for (int j = 0; j < 7; j++) gl[j] = AT(g, j);
    const double f0 = AT(F, 0);
    for (int j = 0; j < m; j++) {
        double t = (AT(F, j + 1) - f0) * k.r[j + 1][0];
        for (int l = 1; l <= j; l++) t = (t - gl[l - 1]) * k.r[j + 1][l];
        gl[j] = t;
        AT(g, j) = t;
    }
    for (int j = 0; j < m; j++) {
        double s = k.c[j][j] * gl[j];
        for (int l = j + 1; l < 7; l++) s = s + k.c[l][j] * gl[l];
        AT(b, j) = s;
    }
}

// Block-wise maxima of |p| and |r| (0 for empty blocks) into part[2 blk + 0/1].
__device__ void block_max2(double mp, double mr, double *part)
{
    __shared__ double sp[IAS_BLOCK], sr[IAS_BLOCK];
    sp[threadIdx.x] = mp;
    sr[threadIdx.x] = mr;
    __syncthreads();
    for (int s = IAS_BLOCK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sp[threadIdx.x] = fmax(sp[threadIdx.x], sp[threadIdx.x + s]);
            sr[threadIdx.x] = fmax(sr[threadIdx.x], sr[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sp[0];
        part[2 * blockIdx.x + 1] = sr[0];
    }
}

// convergence of the corrector: max |b6 - b6_prev| and max |F7|, then the
// previous b takes the current values (src/integrator_ias15.c:419-434)
__global__ __launch_bounds__(IAS_BLOCK) void k_ias_conv(const double *__restrict__ b,
                                                        double *__restrict__ b0,
                                                        const double *__restrict__ F7,
                                                        double *__restrict__ part, size_t n3)
{
    double md = 0.0, ma = 0.0;
    for (size_t q = blockIdx.x * static_cast<size_t>(IAS_BLOCK) + threadIdx.x; q < n3;
         q += static_cast<size_t>(gridDim.x) * IAS_BLOCK) {
        md = fmax(md, fabs(AT(b, 6) - AT(b0, 6)));
        ma = fmax(ma, fabs(F7[q]));
#pragma unroll
        for (int j = 0; j < 7; j++) AT(b0, j) = AT(b, j);
    }
    block_max2(md, ma, part);
}

// max |b6| and max |a1| for the step's error estimate (:455)
__global__ __launch_bounds__(IAS_BLOCK) void k_ias_errmax(const double *__restrict__ b,
                                                          const double *__restrict__ a1,
                                                          double *__restrict__ part, size_t n3)
{
    double mb = 0.0, ma = 0.0;
    for (size_t q = blockIdx.x * static_cast<size_t>(IAS_BLOCK) + threadIdx.x; q < n3;
         q += static_cast<size_t>(gridDim.x) * IAS_BLOCK) {
        mb = fmax(mb, fabs(AT(b, 6)));
        ma = fmax(ma, fabs(a1[q]));
    }
    block_max2(mb, ma, part);
}

// max |p| and max |r| of two plain arrays (initial step size)
__global__ __launch_bounds__(IAS_BLOCK) void k_absmax2(const double *__restrict__ p,
                                                       const double *__restrict__ r,
                                                       double *__restrict__ part, size_t n3)
{
    double mp = 0.0, mr = 0.0;
    for (size_t q = blockIdx.x * static_cast<size_t>(IAS_BLOCK) + threadIdx.x; q < n3;
         q += static_cast<size_t>(gridDim.x) * IAS_BLOCK) {
        mp = fmax(mp, fabs(p[q]));
        mr = fmax(mr, fabs(r[q]));
    }
    block_max2(mp, mr, part);
}

// end point of the step with compensated sums (src/integrator_ias15.c:921-983);
// the compensation terms start from the accepted ones (txc = xc + ...)
__global__ void k_ias_step(double *__restrict__ x1, double *__restrict__ v1,
                           double *__restrict__ txc, double *__restrict__ tvc,
                           const double *__restrict__ x, const double *__restrict__ v,
                           const double *__restrict__ a, const double *__restrict__ b,
                           const double *__restrict__ xc, const double *__restrict__ vc, double dt,
                           size_t n3)
{
    const size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (q >= n3) return;
    const double b0 = AT(b, 0), b1 = AT(b, 1), b2 = AT(b, 2), b3 = AT(b, 3), b4 = AT(b, 4),
                 b5 = AT(b, 5), b6 = AT(b, 6);
    const double sx = a[q] + b0 / 3.0 + b1 / 6.0 + b2 / 10.0 + b3 / 15.0 + b4 / 21.0 + b5 / 28.0 +
                      b6 / 36.0;
    double cx = xc[q];
    cx += dt * (v[q] + dt * sx / 2.0);
    const double xn = x[q] + cx;
    cx += (x[q] - xn);
    x1[q] = xn;
    txc[q] = cx;

    const double sv = a[q] + b0 / 2.0 + b1 / 3.0 + b2 / 4.0 + b3 / 5.0 + b4 / 6.0 + b5 / 7.0 +
                      b6 / 8.0;
    double cv = vc[q];
    cv += dt * sv;
    const double vn = v[q] + cv;
    cv += (v[q] - vn);
    v1[q] = vn;
    tvc[q] = cv;
}

// b for the next step (src/integrator_ias15.c:1276-1382)
__global__ void k_ias_refine(double *__restrict__ b, double *__restrict__ e, double q1, int refine,
                             size_t n3)
{
    const size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (q >= n3) return;
    const double q2 = q1 * q1, q3 = q2 * q1, q4 = q3 * q1, q5 = q4 * q1, q6 = q5 * q1,
                 q7 = q6 * q1;
    double db[7];
#pragma unroll
    for (int j = 0; j < 7; j++) db[j] = refine ? AT(b, j) - AT(e, j) : 0.0;
    const double b0 = AT(b, 0), b1 = AT(b, 1), b2 = AT(b, 2), b3 = AT(b, 3), b4 = AT(b, 4),
                 b5 = AT(b, 5), b6 = AT(b, 6);
    double en[7];
    en[0] = q1 * (b6 * 7.0 + b5 * 6.0 + b4 * 5.0 + b3 * 4.0 + b2 * 3.0 + b1 * 2.0 + b0);
    en[1] = q2 * (b6 * 21.0 + b5 * 15.0 + b4 * 10.0 + b3 * 6.0 + b2 * 3.0 + b1);
    en[2] = q3 * (b6 * 35.0 + b5 * 20.0 + b4 * 10.0 + b3 * 4.0 + b2);
    en[3] = q4 * (b6 * 35.0 + b5 * 15.0 + b4 * 5.0 + b3);
    en[4] = q5 * (b6 * 21.0 + b5 * 6.0 + b4);
    en[5] = q6 * (b6 * 7.0 + b5);
    en[6] = q7 * b6;
#pragma unroll
    for (int j = 0; j < 7; j++) {
        AT(e, j) = en[j];
        AT(b, j) = en[j] + db[j];
    }
}

// x1 = x + dt0 v  /  a1 = a1 - a (initial step size)
__global__ void k_ias_axpy(double *__restrict__ out, const double *__restrict__ x,
                           const double *__restrict__ v, double s, size_t n3)
{
    const size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (q < n3) out[q] = x[q] + s * v[q];
}

__global__ void k_ias_sub(double *__restrict__ a1, const double *__restrict__ a, size_t n3)
{
    const size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (q < n3) a1[q] = a1[q] - a[q];
}

#undef AT

// ---------------------------------------------------------------- host driver
namespace {

struct IasState {
    Context *c = nullptr;
    size_t n3 = 0;
    int grid = 1, rgrid = 1;
    double *mem = nullptr;
    double *x, *v, *m, *a, *x1, *v1, *a1, *F, *g, *b, *b0, *e, *xc, *vc, *txc, *tvc, *part;
    std::vector<double> hpart;

    ErrorStatus alloc(Context *ctx, int n)
    {
        c = ctx;
        n3 = static_cast<size_t>(3) * n;
        grid = static_cast<int>((n3 + IAS_BLOCK - 1) / IAS_BLOCK);
        rgrid = std::min(grid, IAS_MAX_GRID);
        const size_t words = n3 * (10 + 8 + 4 * 7) + n + 2 * IAS_MAX_GRID;
        GS_HIP(hipMalloc(reinterpret_cast<void **>(&mem), words * sizeof(double)));
        GS_HIP(hipMemsetAsync(mem, 0, words * sizeof(double), c->stream));
        double *p = mem;
        auto take = [&](size_t k) {
            double *r = p;
            p += k;
            return r;
        };
        x = take(n3), v = take(n3), a = take(n3), x1 = take(n3), v1 = take(n3), a1 = take(n3);
        xc = take(n3), vc = take(n3), txc = take(n3), tvc = take(n3);
        F = take(8 * n3), g = take(7 * n3), b = take(7 * n3), b0 = take(7 * n3), e = take(7 * n3);
        m = take(static_cast<size_t>(n));
        part = take(2 * IAS_MAX_GRID);
        hpart.resize(2 * IAS_MAX_GRID);
        return make_success_error_status();
    }
    ~IasState()
    {
        if (mem) {
            (void)hipStreamSynchronize(c->stream);
            (void)hipFree(mem);
        }
    }
    // reduce the block maxima written by the last k_* max kernel
    ErrorStatus maxima(double *m0, double *m1)
    {
        GS_HIP(hipMemcpyAsync(hpart.data(), part, sizeof(double) * 2 * rgrid,
                              hipMemcpyDeviceToHost, c->stream));
        GS_HIP(hipStreamSynchronize(c->stream));
        double p0 = 0.0, p1 = 0.0;
        for (int k = 0; k < rgrid; k++) {
            p0 = fmax(p0, hpart[2 * k]);
            p1 = fmax(p1, hpart[2 * k + 1]);
        }
        *m0 = p0;
        *m1 = p1;
        return make_success_error_status();
    }
};

}  // namespace

// src/integrator_ias15.c:702-806
static ErrorStatus ias15_initial_dt(IasState &S, int n, double G, const AccelerationParam *ap,
                                    double *out)
{
    Context *c = S.c;
    hipLaunchKernelGGL(k_absmax2, dim3(S.rgrid), dim3(IAS_BLOCK), 0, c->stream, S.x, S.a, S.part,
                       S.n3);
    GS_HIP(hipGetLastError());
    double d0, d1;
    GS_TRY(S.maxima(&d0, &d1));
    const double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * (d0 / d1);
    hipLaunchKernelGGL(k_ias_axpy, dim3(S.grid), dim3(IAS_BLOCK), 0, c->stream, S.x1, S.x, S.v, dt0,
                       S.n3);
    GS_HIP(hipGetLastError());
    GS_TRY(device_acceleration(c, S.x1, S.m, n, G, ap, S.a1));
    hipLaunchKernelGGL(k_ias_sub, dim3(S.grid), dim3(IAS_BLOCK), 0, c->stream, S.a1, S.a, S.n3);
    hipLaunchKernelGGL(k_absmax2, dim3(S.rgrid), dim3(IAS_BLOCK), 0, c->stream, S.a1, S.a1, S.part,
                       S.n3);
    GS_HIP(hipGetLastError());
    double d2, unused;
    GS_TRY(S.maxima(&d2, &unused));
    d2 = d2 / dt0;
    double dt1;
    if (fmax(d1, d2) <= 1e-15)
        dt1 = fmax(1e-6, dt0 * 1e-3);
    else
        dt1 = pow((0.01 / fmax(d1, d2)), 1.0 / (1.0 + static_cast<double>(15)));
    *out = fmin(100.0 * dt0, dt1);
    if (*out <= 0.0)
        return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR, "Initial dt is less than or equal to zero");
    return make_success_error_status();
}

// The main loop of src/integrator_ias15.c:229-615 on device state.
ErrorStatus ias15_gpu(System *system, IntegratorParam *ip, AccelerationParam *ap,
                      OutputParam *op, SimulationStatus *status, Settings *settings,
                      const double tf)
{
    const int n = system->num_particles;
    if (n <= 0) return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR, "Number of particles must be positive");
    // The corrector iterates to |db6| / |a| < 1e-16 and the step control reads
    // the size of b6: forces with FP32 rounding noise drive both, and the step
    // size, to their limits (dt -> 1e-12 tf), so the FP32 methods are refused.
    if (ap->method == ACCELERATION_METHOD_PAIRWISE_FP32 ||
        ap->method == ACCELERATION_METHOD_MASSLESS_FP32)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR,
                                    "IAS15 needs FP64 forces; acceleration method %d computes in "
                                    "FP32",
                                    ap->method);
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    IasState S;
    GS_TRY(S.alloc(c, n));
    const size_t n3 = S.n3;
    GS_HIP(hipMemcpyAsync(S.x, system->x, sizeof(double) * n3, hipMemcpyHostToDevice, c->stream));
    GS_HIP(hipMemcpyAsync(S.v, system->v, sizeof(double) * n3, hipMemcpyHostToDevice, c->stream));
    GS_HIP(hipMemcpyAsync(S.m, system->m, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    const double G = system->G;
    SnapshotWriter snaps;
    GS_TRY(snaps.begin(c, op, system));

    Ias15Coeffs K;
    ias15_coeffs(&K);
    IasK k;
    for (int i = 0; i < 7; i++)
        for (int j = 0; j < 7; j++) k.c[i][j] = K.c[i][j];
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) k.r[i][j] = K.r[i][j];

    const double tolerance = ip->tolerance;
    const double safety_fac = 0.25, exponent = 1.0 / 7.0, tolerance_pc = 1e-16;
    const dim3 grid(S.grid), block(IAS_BLOCK), rgrid(S.rgrid);

    GS_TRY(device_acceleration(c, S.x, S.m, n, G, ap, S.a));
    double dt;
    if (ip->initial_dt > 0.0) {
        dt = ip->initial_dt;
    } else {
        GS_TRY(ias15_initial_dt(S, n, G, ap, &dt));
        if (dt > tf) dt = tf;
    }

    // initial snapshot after the initial step size, before the status is
    // reset (src/integrator_ias15.c:381-410)
    if (snaps.on() && op->output_initial)
        GS_TRY(snaps.write_host(system->x, system->v, status->t, status->dt));
    status->t = 0.0;
    status->dt = dt;
    status->num_steps = 0;
    bool refine_flag = false;
    while (status->t < tf) {
        for (int it = 0; it < 12; it++) {
            for (int i = 0; i < 8; i++) {
                hipLaunchKernelGGL(k_ias_predict, grid, block, 0, c->stream, S.x1, S.v1, S.x, S.v,
                                   S.a, S.b, S.xc, S.vc, K.h[i], dt, n3);
                GS_HIP(hipGetLastError());
                GS_TRY(device_acceleration(c, S.x1, S.m, n, G, ap, S.F + i * n3));
                hipLaunchKernelGGL(k_ias_gb, grid, block, 0, c->stream, S.g, S.b, S.F, i, k, n3);
                GS_HIP(hipGetLastError());
            }
            hipLaunchKernelGGL(k_ias_conv, rgrid, block, 0, c->stream, S.b, S.b0, S.F + 7 * n3,
                               S.part, n3);
            GS_HIP(hipGetLastError());
            double md, ma;
            GS_TRY(S.maxima(&md, &ma));
            if ((md / ma) < tolerance_pc) break;
        }

        hipLaunchKernelGGL(k_ias_step, grid, block, 0, c->stream, S.x1, S.v1, S.txc, S.tvc, S.x,
                           S.v, S.a, S.b, S.xc, S.vc, dt, n3);
        GS_HIP(hipGetLastError());
        GS_TRY(device_acceleration(c, S.x1, S.m, n, G, ap, S.a1));
        hipLaunchKernelGGL(k_ias_errmax, rgrid, block, 0, c->stream, S.b, S.a1, S.part, n3);
        GS_HIP(hipGetLastError());
        double mb, ma1;
        GS_TRY(S.maxima(&mb, &ma1));
        const double error_b7 = mb / ma1;
        double error = pow((error_b7 / tolerance), exponent);
        if (error < 1e-10) error = 1e-10;
        const double dt_new = dt / error;

        if (error <= 1.0 || dt == tf * 1e-12) {
            status->num_steps++;
            status->t += dt;
            hipLaunchKernelGGL(k_ias_refine, grid, block, 0, c->stream, S.b, S.e, dt_new / dt,
                               refine_flag ? 1 : 0, n3);
            GS_HIP(hipGetLastError());
            refine_flag = true;
            // accepted: the trial state and its compensation terms become current
            std::swap(S.xc, S.txc);
            std::swap(S.vc, S.tvc);
            std::swap(S.x, S.x1);
            std::swap(S.v, S.v1);
            std::swap(S.a, S.a1);
            if (snaps.due(status->t)) GS_TRY(snaps.write_device(S.x, S.v, status->t, dt));
        }

        if (dt_new > (dt / safety_fac))
            dt /= safety_fac;
        else if (dt_new < dt * safety_fac)
            dt *= safety_fac;
        else
            dt = dt_new;
        if (dt_new < tf * 1e-12) dt = tf * 1e-12;
        if ((status->t < tf) && (status->t + dt > tf)) dt = tf - status->t;
        status->dt = dt;
        // a non-finite error estimate would leave dt NaN and the loop without
        // progress (the reference spins there); fail instead
        if (!(dt > 0.0) || !std::isfinite(dt))
            return WRAP_RAISE_ERROR_FMT(GRAV_FAILURE, "IAS15 step size became invalid (%g) at t = %g",
                                        dt, status->t);
        if (settings->is_exit_ptr && *settings->is_exit_ptr) break;
    }

    GS_TRY(snaps.finish());
    GS_HIP(hipMemcpyAsync(system->x, S.x, sizeof(double) * n3, hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipMemcpyAsync(system->v, S.v, sizeof(double) * n3, hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    return make_success_error_status();
}

}  // namespace gs
