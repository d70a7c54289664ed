// direct.hip -- direct-summation force kernels for gfx950.
//
//  k_direct_f64    FP64, one target per lane, sources staged through LDS in
//                  index order: every target sums its sources l = 0..N-1
//                  (l != k) in the order the reference's symmetric i<j loop
//                  produces (src/acceleration.c:198-231), with the same
//                  per-term operation order, contraction off and IEEE sqrt
//                  and divide, so results are bitwise equal to the reference.
//  k_direct_f32    FP32 split precision: sources converted once per target
//                  block to FP32 offsets from the block centre and staged as
//                  float4 tiles in LDS; each lane holds IPT targets in
//                  registers; tile partial sums in plain FP32 are folded into
//                  Kahan-compensated running sums; one rsqrt per interaction.
//  k_reduce_partials  FP64 fixed-order sum of the per-source-chunk partials.
//  k_energy        FP64 kinetic + pair potential energy (src/utils.c:27-59).
//  k_lf_*          Leapfrog updates with the reference's compensated
//                  summation (src/integrator.c:973-1094).
#include "gs_internal.h"

#pragma clang fp contract(off)

namespace gs {

constexpr int F64_BLOCK = 256;

// ------------------------------------------------------------------ FP64
template <bool SKIP_MASSLESS>
__global__ __launch_bounds__(F64_BLOCK) void k_direct_f64(const double *__restrict__ x,
                                                          const double *__restrict__ m, int n,
                                                          double G, double eps2,
                                                          const int *__restrict__ targets,
                                                          int num_targets,
                                                          double *__restrict__ a)
{
    __shared__ double sx[F64_BLOCK], sy[F64_BLOCK], sz[F64_BLOCK], sm[F64_BLOCK];
    const int t = blockIdx.x * F64_BLOCK + threadIdx.x;
    const bool live = t < num_targets;
    const int k = live ? (targets ? targets[t] : t) : -1;
    double xk0 = 0.0, xk1 = 0.0, xk2 = 0.0;
    if (live) {
        xk0 = x[3 * k + 0];
        xk1 = x[3 * k + 1];
        xk2 = x[3 * k + 2];
    }
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int base = 0; base < n; base += F64_BLOCK) {
        const int j = base + threadIdx.x;
        if (j < n) {
            sx[threadIdx.x] = x[3 * j + 0];
            sy[threadIdx.x] = x[3 * j + 1];
            sz[threadIdx.x] = x[3 * j + 2];
            sm[threadIdx.x] = m[j];
        }
        __syncthreads();
        const int cnt = min(F64_BLOCK, n - base);
        if (live) {
            for (int jj = 0; jj < cnt; jj++) {
                const double ml = sm[jj];
                if (base + jj == k) continue;
                if (SKIP_MASSLESS && ml == 0.0) continue;
                const double R0 = sx[jj] - xk0;
                const double R1 = sy[jj] - xk1;
                const double R2 = sz[jj] - xk2;
                const double r = __builtin_sqrt(R0 * R0 + R1 * R1 + R2 * R2 + eps2);
                const double tv = G / (r * r * r);
                a0 += (tv * R0) * ml;
                a1 += (tv * R1) * ml;
                a2 += (tv * R2) * ml;
            }
        }
        __syncthreads();
    }
    if (live) {
        a[3 * t + 0] = a0;
        a[3 * t + 1] = a1;
        a[3 * t + 2] = a2;
    }
}

hipError_t launch_direct_f64(Context *c, const double *x, const double *m, int n, double G,
                             double eps, bool skip_massless, double *a)
{
    const int grid = (n + F64_BLOCK - 1) / F64_BLOCK;
    const int rec = timing_begin(c, "direct_f64");
    if (skip_massless)
        hipLaunchKernelGGL(k_direct_f64<true>, dim3(grid), dim3(F64_BLOCK), 0, c->stream, x, m,
                           n, G, eps * eps, nullptr, n, a);
    else
        hipLaunchKernelGGL(k_direct_f64<false>, dim3(grid), dim3(F64_BLOCK), 0, c->stream, x, m,
                           n, G, eps * eps, nullptr, n, a);
    timing_end(c, rec);
    return hipGetLastError();
}

// FP64 direct sum for a list of targets (the GPU FP64 oracle for sampled
// error checks); out[t] is the acceleration of particle targets[t].
hipError_t launch_direct_f64_targets(Context *c, const double *x, const double *m, int n,
                                     double G, double eps, bool skip_massless,
                                     const int *targets, int num_targets, double *out)
{
    const int grid = (num_targets + F64_BLOCK - 1) / F64_BLOCK;
    if (skip_massless)
        hipLaunchKernelGGL(k_direct_f64<true>, dim3(grid), dim3(F64_BLOCK), 0, c->stream, x, m,
                           n, G, eps * eps, targets, num_targets, out);
    else
        hipLaunchKernelGGL(k_direct_f64<false>, dim3(grid), dim3(F64_BLOCK), 0, c->stream, x, m,
                           n, G, eps * eps, targets, num_targets, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------ FP32 split
// Kahan step: fold tile partial p into (s, comp).  Contraction is off for
// this file, so the compensation survives as written.
__device__ __forceinline__ void kahan_add(float &s, float &comp, float p)
{
    const float y = p - comp;
    const float t = s + y;
    comp = (t - s) - y;
    s = t;
}

template <int BS, int IPT, bool SOFT>
__global__ __launch_bounds__(BS) void k_direct_f32(const double *__restrict__ x,
                                                   const double *__restrict__ m, int n,
                                                   float eps2, int chunk,
                                                   double *__restrict__ partial)
{
    __shared__ float4 tile[2][BS];
    const int tid = threadIdx.x;
    const int i0 = blockIdx.x * (BS * IPT);
    // block centre: the first target of the block (FP64)
    const double c0 = x[3 * i0 + 0], c1 = x[3 * i0 + 1], c2 = x[3 * i0 + 2];

    float xi[IPT], yi[IPT], zi[IPT];
    float px[IPT], py[IPT], pz[IPT];
    float sx[IPT], sy[IPT], sz[IPT];
    float cx[IPT], cy[IPT], cz[IPT];
#pragma unroll
    for (int r = 0; r < IPT; r++) {
        const int i = i0 + r * BS + tid;
        const int ii = i < n ? i : i0;
        xi[r] = static_cast<float>(x[3 * ii + 0] - c0);
        yi[r] = static_cast<float>(x[3 * ii + 1] - c1);
        zi[r] = static_cast<float>(x[3 * ii + 2] - c2);
        px[r] = py[r] = pz[r] = 0.0f;
        sx[r] = sy[r] = sz[r] = 0.0f;
        cx[r] = cy[r] = cz[r] = 0.0f;
    }

    const int jb = blockIdx.y * chunk;
    const int je = min(n, jb + chunk);
    auto stage = [&](int buf, int base) {
        const int j = base + tid;
        float4 s;
        if (j < je) {
            s.x = static_cast<float>(x[3 * j + 0] - c0);
            s.y = static_cast<float>(x[3 * j + 1] - c1);
            s.z = static_cast<float>(x[3 * j + 2] - c2);
            s.w = static_cast<float>(m[j]);
        } else {
            s.x = s.y = s.z = 0.0f;
            s.w = 0.0f;  // zero mass: contributes exactly 0
        }
        tile[buf][tid] = s;
    };

    int buf = 0;
    if (jb < je) stage(0, jb);
    __syncthreads();
    for (int base = jb; base < je; base += BS) {
        if (base + BS < je) stage(buf ^ 1, base + BS);
#pragma unroll 8
        for (int jj = 0; jj < BS; jj++) {
            const float4 s = tile[buf][jj];
#pragma unroll
            for (int r = 0; r < IPT; r++) {
                const float dx = s.x - xi[r];
                const float dy = s.y - yi[r];
                const float dz = s.z - zi[r];
                float r2 = __builtin_fmaf(dx, dx, eps2);
                r2 = __builtin_fmaf(dy, dy, r2);
                r2 = __builtin_fmaf(dz, dz, r2);
                float ri = __builtin_amdgcn_rsqf(r2);
                if (!SOFT) ri = r2 > 0.0f ? ri : 0.0f;
                const float mr = s.w * ri;
                const float w = mr * (ri * ri);
                px[r] = __builtin_fmaf(dx, w, px[r]);
                py[r] = __builtin_fmaf(dy, w, py[r]);
                pz[r] = __builtin_fmaf(dz, w, pz[r]);
            }
        }
#pragma unroll
        for (int r = 0; r < IPT; r++) {
            kahan_add(sx[r], cx[r], px[r]);
            kahan_add(sy[r], cy[r], py[r]);
            kahan_add(sz[r], cz[r], pz[r]);
            px[r] = py[r] = pz[r] = 0.0f;
        }
        __syncthreads();
        buf ^= 1;
    }

    double *out = partial + static_cast<size_t>(blockIdx.y) * 3 * n;
#pragma unroll
    for (int r = 0; r < IPT; r++) {
        const int i = i0 + r * BS + tid;
        if (i < n) {
            out[3 * i + 0] = static_cast<double>(sx[r]) - static_cast<double>(cx[r]);
            out[3 * i + 1] = static_cast<double>(sy[r]) - static_cast<double>(cy[r]);
            out[3 * i + 2] = static_cast<double>(sz[r]) - static_cast<double>(cz[r]);
        }
    }
}

__global__ void k_reduce_partials(const double *__restrict__ partial, int nsplit, int n3,
                                  double G, double *__restrict__ a)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n3) return;
    double s = 0.0;
    for (int k = 0; k < nsplit; k++) s += partial[static_cast<size_t>(k) * n3 + q];
    a[q] = G * s;
}

Fp32Config &fp32_config()
{
    static Fp32Config cfg;
    return cfg;
}

static int resolve_split(int n, const Fp32Config &cfg, int cu_count)
{
    if (cfg.jsplit > 0) return cfg.jsplit;
    const int tb = cfg.block * cfg.ipt;
    const int blocks = (n + tb - 1) / tb;
    const int want = 8 * (cu_count > 0 ? cu_count : 256);  // workgroups to fill the chip
    int split = (want + blocks - 1) / blocks;
    const int max_split = (n + 4 * cfg.block - 1) / (4 * cfg.block);  // >= 4 tiles per chunk
    if (split > max_split) split = max_split;
    return split < 1 ? 1 : split;
}

size_t direct_f32_partial_bytes(int n, const Fp32Config &cfg, int cu_count)
{
    return static_cast<size_t>(resolve_split(n, cfg, cu_count)) * 3 * n * sizeof(double);
}

template <int BS, int IPT>
static void launch_f32_shape(Context *c, dim3 grid, const double *x, const double *m, int n,
                             float eps2, bool soft, int chunk, double *partial)
{
    if (soft)
        hipLaunchKernelGGL((k_direct_f32<BS, IPT, true>), grid, dim3(BS), 0, c->stream, x, m, n,
                           eps2, chunk, partial);
    else
        hipLaunchKernelGGL((k_direct_f32<BS, IPT, false>), grid, dim3(BS), 0, c->stream, x, m,
                           n, eps2, chunk, partial);
}

hipError_t launch_direct_f32(Context *c, const double *x, const double *m, int n, double G,
                             double eps, double *a, void *partial_buf, size_t partial_bytes,
                             const Fp32Config &cfg)
{
    const int split = resolve_split(n, cfg, c->cu_count);
    if (partial_bytes < static_cast<size_t>(split) * 3 * n * sizeof(double))
        return hipErrorInvalidValue;
    const int tb = cfg.block * cfg.ipt;
    const int blocks = (n + tb - 1) / tb;
    int chunk = (n + split - 1) / split;
    chunk = (chunk + cfg.block - 1) / cfg.block * cfg.block;
    const int nsplit = (n + chunk - 1) / chunk;
    dim3 grid(blocks, nsplit);
    const float eps2 = static_cast<float>(eps * eps);
    const bool soft = eps2 > 0.0f;
    double *partial = static_cast<double *>(partial_buf);
    const int rec = timing_begin(c, "direct_f32");
    switch (cfg.block * 16 + cfg.ipt) {
        case 256 * 16 + 1: launch_f32_shape<256, 1>(c, grid, x, m, n, eps2, soft, chunk, partial); break;
        case 256 * 16 + 2: launch_f32_shape<256, 2>(c, grid, x, m, n, eps2, soft, chunk, partial); break;
        case 256 * 16 + 4: launch_f32_shape<256, 4>(c, grid, x, m, n, eps2, soft, chunk, partial); break;
        case 128 * 16 + 2: launch_f32_shape<128, 2>(c, grid, x, m, n, eps2, soft, chunk, partial); break;
        case 128 * 16 + 4: launch_f32_shape<128, 4>(c, grid, x, m, n, eps2, soft, chunk, partial); break;
        case 512 * 16 + 1: launch_f32_shape<512, 1>(c, grid, x, m, n, eps2, soft, chunk, partial); break;
        case 512 * 16 + 2: launch_f32_shape<512, 2>(c, grid, x, m, n, eps2, soft, chunk, partial); break;
        default: timing_end(c, rec); return hipErrorInvalidValue;
    }
    timing_end(c, rec);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int n3 = 3 * n;
    hipLaunchKernelGGL(k_reduce_partials, dim3((n3 + 255) / 256), dim3(256), 0, c->stream,
                       partial, nsplit, n3, G, a);
    return hipGetLastError();
}

// ------------------------------------------------------------------ energy
constexpr int EN_BLOCK = 256;

__global__ __launch_bounds__(EN_BLOCK) void k_energy(const double *__restrict__ x,
                                                     const double *__restrict__ v,
                                                     const double *__restrict__ m, int n,
                                                     double G, double *__restrict__ partials)
{
    __shared__ double red[EN_BLOCK];
    const int i = blockIdx.x * EN_BLOCK + threadIdx.x;
    double e = 0.0;
    if (i < n) {
        const double v0 = v[3 * i], v1 = v[3 * i + 1], v2 = v[3 * i + 2];
        const double vn = __builtin_sqrt(v0 * v0 + v1 * v1 + v2 * v2);
        e = 0.5 * m[i] * vn * vn;
        const double x0 = x[3 * i], x1 = x[3 * i + 1], x2 = x[3 * i + 2];
        const double gm = G * m[i];
        for (int j = i + 1; j < n; j++) {
            const double r0 = x0 - x[3 * j], r1 = x1 - x[3 * j + 1], r2 = x2 - x[3 * j + 2];
            e -= gm * m[j] / __builtin_sqrt(r0 * r0 + r1 * r1 + r2 * r2);
        }
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int s = EN_BLOCK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

hipError_t launch_energy(Context *c, const double *x, const double *v, const double *m, int n,
                         double G, double *partials, int *num_partials)
{
    const int grid = (n + EN_BLOCK - 1) / EN_BLOCK;
    *num_partials = grid;
    hipLaunchKernelGGL(k_energy, dim3(grid), dim3(EN_BLOCK), 0, c->stream, x, v, m, n, G,
                       partials);
    return hipGetLastError();
}

// ------------------------------------------------------------------ leapfrog
__global__ void k_lf_half_kick(double *__restrict__ v, double *__restrict__ ve,
                               const double *__restrict__ a, int n3, double dt)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n3) return;
    const double v0 = v[q];
    double e = ve[q] + 0.5 * a[q] * dt;
    const double v1 = v0 + e;
    e += v0 - v1;
    v[q] = v1;
    ve[q] = e;
}

__global__ void k_lf_drift(double *__restrict__ x, double *__restrict__ xe,
                           const double *__restrict__ v, int n3, double dt)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n3) return;
    const double x0 = x[q];
    double e = xe[q] + v[q] * dt;
    const double x1 = x0 + e;
    e += x0 - x1;
    x[q] = x1;
    xe[q] = e;
}

__global__ void k_lf_kick(double *__restrict__ v, double *__restrict__ ve,
                          const double *__restrict__ a, int n3, double dt)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n3) return;
    const double v0 = v[q];
    double e = ve[q] + a[q] * dt;
    const double v1 = v0 + e;
    e += v0 - v1;
    v[q] = v1;
    ve[q] = e;
}

__global__ void k_lf_sync(double *__restrict__ v, const double *__restrict__ a, int n3,
                          double dt)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n3) return;
    v[q] -= 0.5 * a[q] * dt;
}

static dim3 grid3(int n) { return dim3((3 * n + 255) / 256); }

hipError_t launch_lf_half_kick(Context *c, double *v, double *ve, const double *a, int n,
                               double dt)
{
    hipLaunchKernelGGL(k_lf_half_kick, grid3(n), dim3(256), 0, c->stream, v, ve, a, 3 * n, dt);
    return hipGetLastError();
}

hipError_t launch_lf_drift(Context *c, double *x, double *xe, const double *v, int n, double dt)
{
    hipLaunchKernelGGL(k_lf_drift, grid3(n), dim3(256), 0, c->stream, x, xe, v, 3 * n, dt);
    return hipGetLastError();
}

hipError_t launch_lf_kick(Context *c, double *v, double *ve, const double *a, int n, double dt)
{
    hipLaunchKernelGGL(k_lf_kick, grid3(n), dim3(256), 0, c->stream, v, ve, a, 3 * n, dt);
    return hipGetLastError();
}

hipError_t launch_lf_sync(Context *c, double *v, const double *a, int n, double dt)
{
    hipLaunchKernelGGL(k_lf_sync, grid3(n), dim3(256), 0, c->stream, v, a, 3 * n, dt);
    return hipGetLastError();
}

}  // namespace gs
