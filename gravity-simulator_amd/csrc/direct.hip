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
// Bitwise reference sums: every target adds its terms in source-index order,
// term = fl(fl(G / fl(fl(r*r)*r) * fl(x_l - x_k)) * m_l), l != k
// (src/acceleration.c:206-229 as restated per target).  A workgroup takes
// TB targets; per tile of 256 sources every thread computes its source's
// terms for all TB targets into LDS (the expensive sqrt and division run on
// all lanes), then one thread per (target, component) adds the tile's terms
// in index order.  The additions are the reference's, in the reference's
// order, so the result is bit for bit the serial one.  A source the
// reference skips (the target itself; massless sources in the massless form;
// past n) stores +0.0 instead: the running sum starts at +0.0 and a sum in
// round-to-nearest is -0.0 only when both operands are, so the sum is never
// -0.0 and adding +0.0 leaves it unchanged bit for bit.  The adds then form
// one dependent chain of plain FP64 additions per (target, component), with
// no selects.  TB = 4 for small N (more workgroups to fill the chip), 8
// otherwise.
// With LF the kernel is one whole Leapfrog step of the FP64 direct sum: every
// position it reads (source or target) is first drifted as k_lf_drift drifts
// it (compensated, src/integrator.c:1008-1018), and the thread that holds a
// target component's sum then kicks that velocity component as k_lf_kick
// does (src/integrator.c:1030-1039) and writes the drifted position, so the
// three kernels' arithmetic is unchanged and one launch replaces three.
// Positions, their compensation and velocities are read from one set of
// buffers and written to another (every workgroup reads all positions while
// others write theirs); a velocity's compensation is read and written only
// by its own thread, in place.
struct LfStep {
    const double *xe = nullptr, *v = nullptr;  // state at the start of the step (with x)
    double *xo = nullptr, *xeo = nullptr, *vo = nullptr;  // drifted positions, kicked velocities
    double *ve = nullptr;                                  // velocity compensation
    double dt = 0.0;
};

__device__ __forceinline__ double lf_drifted(double x0, double xe, double v, double dt, double &eo)
{
    double e = xe + v * dt;
    const double x1 = x0 + e;
    e += x0 - x1;
    eo = e;
    return x1;
}

template <int TB, bool SKIP_MASSLESS, bool LF>
__global__ __launch_bounds__(F64_BLOCK) void k_direct_f64(const double *__restrict__ x,
                                                          const double *__restrict__ m, int n,
                                                          double G, double eps2,
                                                          const int *__restrict__ targets,
                                                          int num_targets,
                                                          double *__restrict__ a, LfStep lf)
{
    __shared__ double term[TB][3][F64_BLOCK];
    __shared__ double tx[TB][3];
    __shared__ int tk[TB];
    const int tid = threadIdx.x;
    const int t0 = blockIdx.x * TB;
    auto pos = [&](int q) {
        if (!LF) return x[q];
        double e;
        return lf_drifted(x[q], lf.xe[q], lf.v[q], lf.dt, e);
    };
    if (tid < TB) {
        const int t = t0 + tid;
        const int k = t < num_targets ? (targets ? targets[t] : t) : -1;
        tk[tid] = k;
        tx[tid][0] = k >= 0 ? pos(3 * k + 0) : 0.0;
        tx[tid][1] = k >= 0 ? pos(3 * k + 1) : 0.0;
        tx[tid][2] = k >= 0 ? pos(3 * k + 2) : 0.0;
    }
    __syncthreads();
    // accumulator of (target tid / 3, component tid % 3) for tid < 3 TB
    double acc = 0.0;
    const int at = tid / 3, ac = tid % 3;
    for (int base = 0; base < n; base += F64_BLOCK) {
        const int j = base + tid;
        const bool src = j < n;
        double xj0 = 0.0, xj1 = 0.0, xj2 = 0.0, mj = 0.0;
        if (src) {
            xj0 = pos(3 * j + 0);
            xj1 = pos(3 * j + 1);
            xj2 = pos(3 * j + 2);
            mj = m[j];
        }
#pragma unroll 2
        for (int tt = 0; tt < TB; tt++) {
            const int k = tk[tt];
            const bool sk = !src || k < 0 || j == k || (SKIP_MASSLESS && mj == 0.0);
            double t0v = 0.0, t1v = 0.0, t2v = 0.0;
            if (!sk) {
                const double R0 = xj0 - tx[tt][0];
                const double R1 = xj1 - tx[tt][1];
                const double R2 = xj2 - tx[tt][2];
                const double r = __builtin_sqrt(R0 * R0 + R1 * R1 + R2 * R2 + eps2);
                const double tv = G / (r * r * r);
                t0v = (tv * R0) * mj;
                t1v = (tv * R1) * mj;
                t2v = (tv * R2) * mj;
            }
            term[tt][0][tid] = t0v;
            term[tt][1][tid] = t1v;
            term[tt][2][tid] = t2v;
        }
        __syncthreads();
        if (tid < 3 * TB) {
            const double *tp = term[at][ac];
#pragma unroll 32
            for (int jj = 0; jj < F64_BLOCK; jj++) acc = acc + tp[jj];
        }
        __syncthreads();
    }
    if (tid < 3 * TB && t0 + at < num_targets) {
        const int q = 3 * (t0 + at) + ac;
        a[q] = acc;
        if (LF) {
            double ex;
            lf.xo[q] = lf_drifted(x[q], lf.xe[q], lf.v[q], lf.dt, ex);
            lf.xeo[q] = ex;
            const double v0 = lf.v[q];
            double e = lf.ve[q] + acc * lf.dt;
            const double v1 = v0 + e;
            e += v0 - v1;
            lf.vo[q] = v1;
            lf.ve[q] = e;
        }
    }
}

template <bool SKIP_MASSLESS, bool LF = false>
static void launch_f64_shape(Context *c, int n, const double *x, const double *m, double G,
                             double eps2, const int *targets, int num_targets, double *out,
                             const LfStep &lf = LfStep())
{
    // 4 targets per workgroup while 8 would leave CUs idle
    if ((num_targets + 7) / 8 < 2 * c->cu_count)
        hipLaunchKernelGGL((k_direct_f64<4, SKIP_MASSLESS, LF>), dim3((num_targets + 3) / 4),
                           dim3(F64_BLOCK), 0, c->stream, x, m, n, G, eps2, targets, num_targets,
                           out, lf);
    else
        hipLaunchKernelGGL((k_direct_f64<8, SKIP_MASSLESS, LF>), dim3((num_targets + 7) / 8),
                           dim3(F64_BLOCK), 0, c->stream, x, m, n, G, eps2, targets, num_targets,
                           out, lf);
}

hipError_t launch_lf_step_f64(Context *c, const double *x, const double *xe, const double *v,
                              const double *m, int n, double G, double eps, bool skip_massless,
                              double dt, double *a, double *xo, double *xeo, double *vo, double *ve)
{
    LfStep lf;
    lf.xe = xe;
    lf.v = v;
    lf.xo = xo;
    lf.xeo = xeo;
    lf.vo = vo;
    lf.ve = ve;
    lf.dt = dt;
    const int rec = timing_begin(c, "direct_f64");
    if (skip_massless)
        launch_f64_shape<true, true>(c, n, x, m, G, eps * eps, nullptr, n, a, lf);
    else
        launch_f64_shape<false, true>(c, n, x, m, G, eps * eps, nullptr, n, a, lf);
    timing_end(c, rec);
    return hipGetLastError();
}

hipError_t launch_direct_f64(Context *c, const double *x, const double *m, int n, double G,
                             double eps, bool skip_massless, double *a)
{
    const int rec = timing_begin(c, "direct_f64");
    if (skip_massless)
        launch_f64_shape<true>(c, n, x, m, G, eps * eps, nullptr, n, a);
    else
        launch_f64_shape<false>(c, n, x, m, G, eps * eps, nullptr, n, a);
    timing_end(c, rec);
    return hipGetLastError();
}

// FP64 direct sum for a list of targets (the GPU FP64 oracle for sampled
// error checks); out[t] is the acceleration of particle targets[t].
hipError_t launch_direct_f64_targets(Context *c, const double *x, const double *m, int n,
                                     double G, double eps, bool skip_massless,
                                     const int *targets, int num_targets, double *out)
{
    if (skip_massless)
        launch_f64_shape<true>(c, n, x, m, G, eps * eps, targets, num_targets, out);
    else
        launch_f64_shape<false>(c, n, x, m, G, eps * eps, targets, num_targets, out);
    return hipGetLastError();
}

// FP64 sums of one source chunk [blockIdx.y * chunk, +chunk) for a list of
// targets, the same per-term arithmetic as k_direct_f64; partial sums per
// chunk (without G) for k_reduce_partials.  Used for the sampled error
// reference, where a few thousand targets alone would occupy a handful of
// CUs for the whole source range.
template <bool SKIP_MASSLESS>
__global__ __launch_bounds__(F64_BLOCK) void k_direct_f64_part(
    const double *__restrict__ x, const double *__restrict__ m, int n, double eps2,
    const int *__restrict__ targets, int num_targets, int chunk, double *__restrict__ partial)
{
    __shared__ double sx[F64_BLOCK], sy[F64_BLOCK], sz[F64_BLOCK], sm[F64_BLOCK];
    const int t = blockIdx.x * F64_BLOCK + threadIdx.x;
    const bool live = t < num_targets;
    const int k = live ? targets[t] : -1;
    double xk0 = 0.0, xk1 = 0.0, xk2 = 0.0;
    if (live) {
        xk0 = x[3 * k + 0];
        xk1 = x[3 * k + 1];
        xk2 = x[3 * k + 2];
    }
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    const int j0 = blockIdx.y * chunk, j1 = min(n, j0 + chunk);
    for (int base = j0; base < j1; base += F64_BLOCK) {
        const int j = base + threadIdx.x;
        if (j < j1) {
            sx[threadIdx.x] = x[3 * j + 0];
            sy[threadIdx.x] = x[3 * j + 1];
            sz[threadIdx.x] = x[3 * j + 2];
            sm[threadIdx.x] = m[j];
        }
        __syncthreads();
        const int cnt = min(F64_BLOCK, j1 - base);
        if (live) {
            for (int jj = 0; jj < cnt; jj++) {
                const double ml = sm[jj];
                if (base + jj == k) continue;
                if (SKIP_MASSLESS && ml == 0.0) continue;
                const double R0 = sx[jj] - xk0;
                const double R1 = sy[jj] - xk1;
                const double R2 = sz[jj] - xk2;
                const double r = __builtin_sqrt(R0 * R0 + R1 * R1 + R2 * R2 + eps2);
                const double tv = 1.0 / (r * r * r);
                a0 += (tv * R0) * ml;
                a1 += (tv * R1) * ml;
                a2 += (tv * R2) * ml;
            }
        }
        __syncthreads();
    }
    if (live) {
        double *o = partial + (static_cast<size_t>(blockIdx.y) * num_targets + t) * 3;
        o[0] = a0;
        o[1] = a1;
        o[2] = a2;
    }
}

__global__ void k_reduce_partials(const double *__restrict__ partial, int nsplit, int n3,
                                  double G, double *__restrict__ a);

static int f64_part_split(int n, int nt)
{
    const int wgs = (nt + F64_BLOCK - 1) / F64_BLOCK;
    int split = (2048 + wgs - 1) / wgs;
    const int max_split = (n + F64_BLOCK - 1) / F64_BLOCK;  // at least one source tile each
    if (split > max_split) split = max_split;
    return split < 1 ? 1 : split;
}

size_t direct_f64_sampled_bytes(int n, int nt)
{
    return static_cast<size_t>(f64_part_split(n, nt)) * 3 * static_cast<size_t>(nt) * sizeof(double);
}

hipError_t launch_direct_f64_sampled(Context *c, const double *x, const double *m, int n, double G,
                                     double eps, bool skip_massless, const int *targets, int nt,
                                     double *out, void *buf, size_t bytes)
{
    if (nt <= 0) return hipSuccess;
    if (bytes < direct_f64_sampled_bytes(n, nt)) return hipErrorInvalidValue;
    const int split = f64_part_split(n, nt);
    const int chunk = (n + split - 1) / split;
    const int nsplit = (n + chunk - 1) / chunk;
    double *partial = static_cast<double *>(buf);
    dim3 grid((nt + F64_BLOCK - 1) / F64_BLOCK, nsplit);
    if (skip_massless)
        hipLaunchKernelGGL(k_direct_f64_part<true>, grid, dim3(F64_BLOCK), 0, c->stream, x, m, n,
                           eps * eps, targets, nt, chunk, partial);
    else
        hipLaunchKernelGGL(k_direct_f64_part<false>, grid, dim3(F64_BLOCK), 0, c->stream, x, m, n,
                           eps * eps, targets, nt, chunk, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_reduce_partials, dim3((3 * nt + 255) / 256), dim3(256), 0, c->stream,
                       partial, nsplit, 3 * nt, G, out);
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

// Scalar-source variant: sources are pre-converted once per evaluation to
// FP32 offsets from the centre of their own source tile (k_f32_sources);
// every lane converts its FP64 target position to the same tile frame once
// per tile, then streams the tile's sources through wave-uniform (scalar)
// loads, so the sources sit in SGPRs and no LDS or barrier is involved.
constexpr int ST_TILE = 256;

__global__ void k_f32_sources(const double *__restrict__ x, const double *__restrict__ m, int n,
                              int n_pad, float4 *__restrict__ S, double4 *__restrict__ C)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_pad) return;
    const int c0 = (j / ST_TILE) * ST_TILE;  // < n for every tile of n_pad
    const double cx = x[3 * c0 + 0], cy = x[3 * c0 + 1], cz = x[3 * c0 + 2];
    if (j == c0) C[j / ST_TILE] = make_double4(cx, cy, cz, 0.0);
    float4 s;
    if (j < n) {
        s.x = static_cast<float>(x[3 * j + 0] - cx);
        s.y = static_cast<float>(x[3 * j + 1] - cy);
        s.z = static_cast<float>(x[3 * j + 2] - cz);
        s.w = static_cast<float>(m[j]);
    } else {
        s.x = s.y = s.z = s.w = 0.0f;
    }
    S[j] = s;
}

template <int IPT, bool SOFT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(IPT <= 2 ? 8 : 4))) void k_direct_f32s(const double *__restrict__ x, int n,
                                                     const float4 *__restrict__ S,
                                                     const double4 *__restrict__ C, float eps2,
                                                     int tiles_per_chunk, int ntiles,
                                                     double *__restrict__ partial)
{
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int base = wave * 64 * IPT;
    if (base >= n) return;
    double xd[IPT], yd[IPT], zd[IPT];
    float sx[IPT], sy[IPT], sz[IPT], cx[IPT], cy[IPT], cz[IPT];
#pragma unroll
    for (int r = 0; r < IPT; r++) {
        const int i = base + r * 64 + lane;
        const int ii = i < n ? i : base;
        xd[r] = x[3 * ii + 0];
        yd[r] = x[3 * ii + 1];
        zd[r] = x[3 * ii + 2];
        sx[r] = sy[r] = sz[r] = 0.0f;
        cx[r] = cy[r] = cz[r] = 0.0f;
    }
    const int J0 = blockIdx.y * tiles_per_chunk;
    const int J1 = min(ntiles, J0 + tiles_per_chunk);
    for (int J = J0; J < J1; J++) {
        const double4 c = C[J];
        float xi[IPT], yi[IPT], zi[IPT], px[IPT], py[IPT], pz[IPT];
#pragma unroll
        for (int r = 0; r < IPT; r++) {
            xi[r] = static_cast<float>(xd[r] - c.x);
            yi[r] = static_cast<float>(yd[r] - c.y);
            zi[r] = static_cast<float>(zd[r] - c.z);
            px[r] = py[r] = pz[r] = 0.0f;
        }
        const float4 *__restrict__ Sj = S + static_cast<size_t>(J) * ST_TILE;
        // unrolled to 16 sources per step with one target per lane; with
        // several targets per lane 4 sources per step keep the registers
        // within 8 waves per SIMD
        constexpr int UNROLL = IPT == 1 ? 16 : 4;
#pragma unroll UNROLL
        for (int jj = 0; jj < ST_TILE; jj++) {
            const float4 s = Sj[jj];
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
        }
    }
    double *out = partial + static_cast<size_t>(blockIdx.y) * 3 * n;
#pragma unroll
    for (int r = 0; r < IPT; r++) {
        const int i = base + r * 64 + lane;
        if (i < n) {
            out[3 * i + 0] = static_cast<double>(sx[r]) - static_cast<double>(cx[r]);
            out[3 * i + 1] = static_cast<double>(sy[r]) - static_cast<double>(cy[r]);
            out[3 * i + 2] = static_cast<double>(sz[r]) - static_cast<double>(cz[r]);
        }
    }
}

// The scalar-source kernel with a shared target group and a software-
// pipelined source stream.  The W wavefronts of a workgroup hold the same 64
// targets and split the chunk's source tiles into W contiguous runs; at the
// end wavefronts 1..W-1 leave their compensated sums in LDS and wavefront 0
// adds the W sums in wavefront order (FP64, a fixed order), so the FP64
// partial sums written per chunk are 1/W of k_direct_f32s's at the same
// number of wavefronts.  Each wavefront reads its run of float4
// sources in batches of PF_BATCH through wave-uniform loads and issues the
// loads of batch b + 1 before the arithmetic of batch b; the last batch
// re-reads its own sources instead of reading past the run.  Arithmetic, tile
// frames and Kahan folds are k_direct_f32s's.
constexpr int PF_BATCH = 4;

template <bool SOFT>
__device__ __forceinline__ void f32_batch(const float4 (&s)[PF_BATCH], float xi, float yi, float zi,
                                          float eps2, float &px, float &py, float &pz)
{
#pragma unroll
    for (int q = 0; q < PF_BATCH; q++) {
        const float dx = s[q].x - xi;
        const float dy = s[q].y - yi;
        const float dz = s[q].z - zi;
        float r2 = __builtin_fmaf(dx, dx, eps2);
        r2 = __builtin_fmaf(dy, dy, r2);
        r2 = __builtin_fmaf(dz, dz, r2);
        float ri = __builtin_amdgcn_rsqf(r2);
        if (!SOFT) ri = r2 > 0.0f ? ri : 0.0f;
        const float mr = s[q].w * ri;
        const float w = mr * (ri * ri);
        px = __builtin_fmaf(dx, w, px);
        py = __builtin_fmaf(dy, w, py);
        pz = __builtin_fmaf(dz, w, pz);
    }
}

template <int W, bool SOFT>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(8))) void k_direct_f32p(
    const double *__restrict__ x, int n, const float4 *__restrict__ S,
    const double4 *__restrict__ C, float eps2, int tiles_per_chunk, int ntiles,
    double *__restrict__ partial)
{
    static_assert(ST_TILE % (2 * PF_BATCH) == 0, "a tile must hold whole batch pairs");
    __shared__ double other[W - 1][3][64];
    const int lane = threadIdx.x & 63;
    // wave-uniform by construction; readfirstlane lets the compiler see it,
    // so the source addresses below stay scalar
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i0 = blockIdx.x * 64;  // the workgroup's 64 targets (uniform)
    if (i0 >= n) return;
    const int i = i0 + lane;
    const int ii = i < n ? i : i0;
    const double xd = x[3 * ii + 0], yd = x[3 * ii + 1], zd = x[3 * ii + 2];
    float sx = 0.0f, sy = 0.0f, sz = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
    const int J0 = blockIdx.y * tiles_per_chunk;
    const int Jc = min(ntiles, J0 + tiles_per_chunk);
    const int tpw = (Jc - J0 + W - 1) / W;
    const int Ja = min(Jc, J0 + wv * tpw);
    const int Jb = min(Jc, Ja + tpw);
    if (Ja < Jb) {
        const float4 *__restrict__ base = S + static_cast<size_t>(Ja) * ST_TILE;
        const int last = (Jb - Ja) * ST_TILE - PF_BATCH;  // start of the run's last batch
        float4 a[PF_BATCH], b[PF_BATCH];
#pragma unroll
        for (int q = 0; q < PF_BATCH; q++) a[q] = base[q];
        int pos = 0;
        for (int J = Ja; J < Jb; J++) {
            const double4 c = C[J];
            const float xi = static_cast<float>(xd - c.x);
            const float yi = static_cast<float>(yd - c.y);
            const float zi = static_cast<float>(zd - c.z);
            float px = 0.0f, py = 0.0f, pz = 0.0f;
            for (int jj = 0; jj < ST_TILE; jj += 2 * PF_BATCH, pos += 2 * PF_BATCH) {
                const float4 *nb = base + min(pos + PF_BATCH, last);
#pragma unroll
                for (int q = 0; q < PF_BATCH; q++) b[q] = nb[q];
                f32_batch<SOFT>(a, xi, yi, zi, eps2, px, py, pz);
                const float4 *na = base + min(pos + 2 * PF_BATCH, last);
#pragma unroll
                for (int q = 0; q < PF_BATCH; q++) a[q] = na[q];
                f32_batch<SOFT>(b, xi, yi, zi, eps2, px, py, pz);
            }
            kahan_add(sx, cx, px);
            kahan_add(sy, cy, py);
            kahan_add(sz, cz, pz);
        }
    }
    double rx = static_cast<double>(sx) - static_cast<double>(cx);
    double ry = static_cast<double>(sy) - static_cast<double>(cy);
    double rz = static_cast<double>(sz) - static_cast<double>(cz);
    if (wv > 0) {
        other[wv - 1][0][lane] = rx;
        other[wv - 1][1][lane] = ry;
        other[wv - 1][2][lane] = rz;
    }
    __syncthreads();
    if (wv == 0 && i < n) {
#pragma unroll
        for (int w = 0; w < W - 1; w++) {
            rx += other[w][0][lane];
            ry += other[w][1][lane];
            rz += other[w][2][lane];
        }
        double *out = partial + static_cast<size_t>(blockIdx.y) * 3 * n;
        out[3 * i + 0] = rx;
        out[3 * i + 1] = ry;
        out[3 * i + 2] = rz;
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

// Source chunks per evaluation when the configuration leaves it to us:
// about 1024 sources per chunk (128 chunks at N = 2^17).  Measured on
// MI355X at N = 2^17 with 256-source tiles (kernel only): 2048 per chunk
// 4.36 ms, 1024 4.35 ms, 512 4.34 ms; each halving doubles the FP64 partial
// sums (402 MB at 1024) that the reduction reads back (~0.09 ms).
//
// The shared-target kernel (kind 2) spreads a chunk over its W wavefronts:
// 1024 W sources per chunk keep 1024 sources per wavefront (W = 8: 16 chunks
// at N = 2^17, 50 MB of partial sums).
static int resolve_split(int n, const Fp32Config &cfg, int cu_count)
{
    (void)cu_count;
    if (cfg.jsplit > 0) return cfg.jsplit;
    const int per_chunk = cfg.kind == 2 ? 1024 * cfg.block : 1024;
    const int split = (n + per_chunk - 1) / per_chunk;
    return split < 1 ? 1 : split;
}

static int f32s_tiles(int n) { return (n + ST_TILE - 1) / ST_TILE; }

size_t direct_f32_partial_bytes(int n, const Fp32Config &cfg, int cu_count)
{
    size_t b = static_cast<size_t>(resolve_split(n, cfg, cu_count)) * 3 * n * sizeof(double);
    if (cfg.kind >= 1) {
        const size_t nt = f32s_tiles(n);
        b += 256 + nt * ST_TILE * sizeof(float4) + nt * sizeof(double4);
    }
    return b;
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

template <int IPT>
static void launch_f32s_shape(Context *c, dim3 grid, const double *x, int n, const float4 *S,
                              const double4 *C, float eps2, bool soft, int tpc, int ntiles,
                              double *partial)
{
    if (soft)
        hipLaunchKernelGGL((k_direct_f32s<IPT, true>), grid, dim3(256), 0, c->stream, x, n, S, C,
                           eps2, tpc, ntiles, partial);
    else
        hipLaunchKernelGGL((k_direct_f32s<IPT, false>), grid, dim3(256), 0, c->stream, x, n, S,
                           C, eps2, tpc, ntiles, partial);
}

template <int W>
static void launch_f32p_shape(Context *c, dim3 grid, const double *x, int n, const float4 *S,
                              const double4 *C, float eps2, bool soft, int tpc, int ntiles,
                              double *partial)
{
    if (soft)
        hipLaunchKernelGGL((k_direct_f32p<W, true>), grid, dim3(64 * W), 0, c->stream, x, n, S, C,
                           eps2, tpc, ntiles, partial);
    else
        hipLaunchKernelGGL((k_direct_f32p<W, false>), grid, dim3(64 * W), 0, c->stream, x, n, S,
                           C, eps2, tpc, ntiles, partial);
}

hipError_t launch_direct_f32(Context *c, const double *x, const double *m, int n, double G,
                             double eps, double *a, void *partial_buf, size_t partial_bytes,
                             const Fp32Config &cfg)
{
    const int split = resolve_split(n, cfg, c->cu_count);
    if (partial_bytes < direct_f32_partial_bytes(n, cfg, c->cu_count)) return hipErrorInvalidValue;
    if (cfg.kind == 2 && cfg.ipt != 1) return hipErrorInvalidValue;
    const float eps2 = static_cast<float>(eps * eps);
    const bool soft = eps2 > 0.0f;
    double *partial = static_cast<double *>(partial_buf);
    int nsplit;
    if (cfg.kind >= 1) {
        const int ntiles = f32s_tiles(n);
        const int n_pad = ntiles * ST_TILE;
        char *tail = static_cast<char *>(partial_buf) + static_cast<size_t>(split) * 3 * n * sizeof(double);
        tail = reinterpret_cast<char *>((reinterpret_cast<uintptr_t>(tail) + 255) & ~uintptr_t(255));
        float4 *S = reinterpret_cast<float4 *>(tail);
        double4 *C = reinterpret_cast<double4 *>(tail + static_cast<size_t>(n_pad) * sizeof(float4));
        hipLaunchKernelGGL(k_f32_sources, dim3((n_pad + 255) / 256), dim3(256), 0, c->stream, x, m,
                           n, n_pad, S, C);
        const int tpc = (ntiles + split - 1) / split;
        nsplit = (ntiles + tpc - 1) / tpc;
        const int rec = timing_begin(c, "direct_f32");
        if (cfg.kind == 2) {
            dim3 grid((n + 63) / 64, nsplit);
            switch (cfg.block) {
                case 4: launch_f32p_shape<4>(c, grid, x, n, S, C, eps2, soft, tpc, ntiles, partial); break;
                case 8: launch_f32p_shape<8>(c, grid, x, n, S, C, eps2, soft, tpc, ntiles, partial); break;
                case 16: launch_f32p_shape<16>(c, grid, x, n, S, C, eps2, soft, tpc, ntiles, partial); break;
                default: timing_end(c, rec); return hipErrorInvalidValue;
            }
        } else {
            const int waves = (n + 64 * cfg.ipt - 1) / (64 * cfg.ipt);
            dim3 grid((waves + 3) / 4, nsplit);
            switch (cfg.ipt) {
                case 1: launch_f32s_shape<1>(c, grid, x, n, S, C, eps2, soft, tpc, ntiles, partial); break;
                case 2: launch_f32s_shape<2>(c, grid, x, n, S, C, eps2, soft, tpc, ntiles, partial); break;
                case 4: launch_f32s_shape<4>(c, grid, x, n, S, C, eps2, soft, tpc, ntiles, partial); break;
                default: timing_end(c, rec); return hipErrorInvalidValue;
            }
        }
        timing_end(c, rec);
    } else {
        const int tb = cfg.block * cfg.ipt;
        const int blocks = (n + tb - 1) / tb;
        int chunk = (n + split - 1) / split;
        chunk = (chunk + cfg.block - 1) / cfg.block * cfg.block;
        nsplit = (n + chunk - 1) / chunk;
        dim3 grid(blocks, nsplit);
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
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int n3 = 3 * n;
    hipLaunchKernelGGL(k_reduce_partials, dim3((n3 + 255) / 256), dim3(256), 0, c->stream,
                       partial, nsplit, n3, G, a);
    return hipGetLastError();
}

// FP32 direct sum for a list of targets against all n sources (the learned
// correction's labels and spot check).  The shared-target kernel of
// launch_direct_f32 (kind 2, 16 wavefronts per 64 targets): the targets'
// positions are gathered into a compact array and the source range is split
// so that even a few hundred targets fill the chip.
__global__ void k_gather_targets(const double *__restrict__ x, const int *__restrict__ idx,
                                 int nt, double *__restrict__ xt)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 3 * nt) return;
    xt[q] = x[3 * static_cast<size_t>(idx[q / 3]) + q % 3];
}

hipError_t launch_gather_rows3(Context *c, const double *src, const int *idx, int nt, double *out)
{
    if (nt <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_targets, dim3((3 * nt + 255) / 256), dim3(256), 0, c->stream, src,
                       idx, nt, out);
    return hipGetLastError();
}

constexpr int TARGET_WAVES = 16;  // wavefronts per 64 targets in launch_direct_f32_targets

// Source chunks: about 1024 workgroups in all, each wavefront of a chunk's
// workgroup with at least one tile.
static int targets_split(int n, int nt)
{
    const int ntiles = f32s_tiles(n);
    const int wgs = (nt + 63) / 64;
    int split = (1024 + wgs - 1) / wgs;
    if (split > ntiles / TARGET_WAVES) split = ntiles / TARGET_WAVES;
    return split < 1 ? 1 : split;
}

size_t direct_f32_targets_bytes(int n, int nt)
{
    const size_t nt3 = 3 * static_cast<size_t>(nt);
    const size_t nt_tiles = f32s_tiles(n);
    return 512 + nt3 * sizeof(double) * (1 + targets_split(n, nt)) + nt_tiles * ST_TILE * sizeof(float4) +
           nt_tiles * sizeof(double4);
}

hipError_t launch_direct_f32_targets(Context *c, const double *x, const double *m, int n, double G,
                                     double eps, const int *targets, int nt, double *out,
                                     void *buf, size_t bytes)
{
    if (nt <= 0) return hipSuccess;
    if (bytes < direct_f32_targets_bytes(n, nt)) return hipErrorInvalidValue;
    const int ntiles = f32s_tiles(n);
    const int n_pad = ntiles * ST_TILE;
    const int split = targets_split(n, nt);
    const int tpc = (ntiles + split - 1) / split;
    const int nsplit = (ntiles + tpc - 1) / tpc;
    const size_t nt3 = 3 * static_cast<size_t>(nt);
    auto align = [](char *p) {
        return reinterpret_cast<char *>((reinterpret_cast<uintptr_t>(p) + 255) & ~uintptr_t(255));
    };
    char *p = align(static_cast<char *>(buf));
    double *xt = reinterpret_cast<double *>(p);
    double *partial = xt + nt3;
    char *tail = align(reinterpret_cast<char *>(partial + static_cast<size_t>(nsplit) * nt3));
    float4 *S = reinterpret_cast<float4 *>(tail);
    double4 *C = reinterpret_cast<double4 *>(tail + static_cast<size_t>(n_pad) * sizeof(float4));
    hipLaunchKernelGGL(k_gather_targets, dim3((3 * nt + 255) / 256), dim3(256), 0, c->stream, x,
                       targets, nt, xt);
    hipLaunchKernelGGL(k_f32_sources, dim3((n_pad + 255) / 256), dim3(256), 0, c->stream, x, m, n,
                       n_pad, S, C);
    const float eps2 = static_cast<float>(eps * eps);
    dim3 grid((nt + 63) / 64, nsplit);
    launch_f32p_shape<TARGET_WAVES>(c, grid, xt, nt, S, C, eps2, eps2 > 0.0f, tpc, ntiles, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_reduce_partials, dim3((3 * nt + 255) / 256), dim3(256), 0, c->stream,
                       partial, nsplit, 3 * nt, G, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------ issue rates
// Chains of one FP32 instruction form (8 independent chains per lane, VGPR
// operands, so the VOP2 encodings v_fmac_f32 / v_mul_f32 and v_rsq_f32):
// their time gives the cycles per wave64 instruction of each form, from
// which an FP32 kernel's instruction mix bounds its rate.  Each thread
// writes one float.
constexpr int ISSUE_ITERS = 4096;

template <int KIND>
__global__ __launch_bounds__(256) void k_issue(float *__restrict__ out, float a0, float b0)
{
    const float a = a0 + threadIdx.x * 1e-9f, b = b0 - threadIdx.x * 1e-9f;
    float x[8];
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = static_cast<float>(threadIdx.x + k + 1);
    for (int i = 0; i < ISSUE_ITERS; i++) {
#pragma unroll
        for (int k = 0; k < 8; k++) {
            if (KIND == 0) x[k] = __builtin_fmaf(a, b, x[k]);
            if (KIND == 1) x[k] = x[k] * a;
            if (KIND == 2) x[k] = __builtin_amdgcn_rsqf(x[k]);
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; k++) s += x[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

hipError_t launch_issue_probe(Context *c, int kind, float *out, int blocks)
{
    switch (kind) {
        case 0: hipLaunchKernelGGL(k_issue<0>, dim3(blocks), dim3(256), 0, c->stream, out, 0.999f, 0.001f); break;
        case 1: hipLaunchKernelGGL(k_issue<1>, dim3(blocks), dim3(256), 0, c->stream, out, 0.999f, 0.001f); break;
        case 2: hipLaunchKernelGGL(k_issue<2>, dim3(blocks), dim3(256), 0, c->stream, out, 0.999f, 0.001f); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

double issue_probe_instructions(int blocks)
{
    return static_cast<double>(blocks) * 256.0 / 64.0 * 8.0 * ISSUE_ITERS;
}

// ------------------------------------------------------------------ copy
// Plain streaming copy (16 B per lane per iteration), used to calibrate the
// HBM byte counters and to measure achievable copy bandwidth.
__global__ void k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, size_t n16)
{
    for (size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; q < n16;
         q += static_cast<size_t>(gridDim.x) * blockDim.x)
        dst[q] = src[q];
}

hipError_t launch_copy(Context *c, const void *src, void *dst, size_t bytes)
{
    const size_t n16 = bytes / 16;
    const int grid = 8 * (c->cu_count > 0 ? c->cu_count : 256);
    const int rec = timing_begin(c, "copy");
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, c->stream,
                       static_cast<const double2 *>(src), static_cast<double2 *>(dst), n16);
    timing_end(c, rec);
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
