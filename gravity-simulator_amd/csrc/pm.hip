// pm.hip -- particle-mesh gravity (src/acceleration_PM.c:15-391), the
// cosmological Leapfrog (src/integrator_cosmology_leapfrog.c:27-255) and the
// cosmological-system helpers (src/system.c:119-250,1337-1380,
// src/cosmology.c:5-26, src/grav_sim.c:140-260) on the GPU.
//
// One force evaluation is a chain of launches on the context stream:
//   k_pm_cell_keys     base cell of each particle (wrapped into the grid)
//   radix sort         (cell, particle) pairs, stable, 8 bits per pass
//   k_pm_cell_ranges   first and end sorted position of each cell's particles
//   k_cic_gather       cloud in cell as a gather: each cell sums, corner by
//                      corner, the particles whose base cell is that corner
//                      away -- no atomics, and a fixed order of the sums
//   hipfftExecD2Z      real -> half spectrum, n x n x (n/2 + 1)
//   k_pm_green         * (-1/k^2) * 3/2 Omega_m H0^2 / (a n^3) / (cell volume * mean density)
//   hipfftExecZ2D      half spectrum -> potential (unnormalised, like FFTW's)
//   k_cic_grad_interp  per particle: +grad phi at its 8 CIC nodes by 4th-order
//                      central differences, then the CIC gather
// The reference's density-contrast pass, delta = rho / (cell volume * mean
// density) - 1, is folded into the Green's function: the transform is
// linear, and the -1 only touches k = 0, which the Green's function zeroes.
// The reference's gradient grid (3 n^3 doubles, written once and gathered
// from) is not formed: each particle evaluates the same stencil expression
// at its own 8 nodes, which gives the same doubles.  Each cell's density is
// summed in one fixed order (corners in the reference's order, particles by
// index within a cell), not the reference's particle-by-particle order (its
// OpenMP loop adds without synchronisation, so it has no fixed order), and
// the scale factors are applied in the spectrum instead of before and after
// the transforms: results agree with the reference's to rounding.
#include <hipfft/hipfft.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "gs_internal.h"
#include "octree.h"
#include "scan.h"

namespace gs {

namespace {

constexpr int PM_BLOCK = 256;

inline dim3 grid_for(size_t work) { return dim3(static_cast<unsigned>((work + PM_BLOCK - 1) / PM_BLOCK)); }

// Periodic boundary conditions, the reference's fmod / parity formula
// (src/system.c:1337-1380).
__device__ __forceinline__ double wrap_coord(double xv, double c, double w)
{
    const double p = xv - c;
    if (p < -w) {
        const double f = fmod(p, w);
        return ((static_cast<int>(p / w) % 2 != 0) ? f + w : f) + c;
    }
    if (p > w) {
        const double f = fmod(p, w);
        return ((static_cast<int>(p / w) % 2 != 0) ? f - w : f) + c;
    }
    return xv;
}

__global__ void k_pm_wrap(double *__restrict__ x, size_t n3, double c0, double c1, double c2,
                          double w)
{
    const size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (q >= n3) return;
    const int d = static_cast<int>(q % 3);
    x[q] = wrap_coord(x[q], d == 0 ? c0 : (d == 1 ? c1 : c2), w);
}

struct Cell {
    int i[3];
    double w[3], wm[3];
};

// Node and weights of a position (src/acceleration_PM.c:44-56).
__device__ __forceinline__ Cell cell_of(const double *xi, double c0, double c1, double c2,
                                        double bw, double inv)
{
    Cell r;
    const double c[3] = {c0, c1, c2};
    for (int d = 0; d < 3; d++) {
        const double xn = ((xi[d] - c[d]) + bw) * inv;
        // (the hardware conversion saturates and maps NaN to 0; the bound
        // keeps i + 1 from overflowing for such inputs)
        r.i[d] = static_cast<int>(fmax(fmin(xn, 1.0e9), -1.0e9));
        r.w[d] = xn - r.i[d];
        r.wm[d] = 1.0 - r.w[d];
    }
    return r;
}

// corner k of the reference's order: (0,0,0) (1,0,0) (0,1,0) (0,0,1) (1,1,0)
// (1,0,1) (0,1,1) (1,1,1)
__device__ __forceinline__ void corner(int k, int &dx, int &dy, int &dz)
{
    dx = (k == 1 || k == 4 || k == 5 || k == 7) ? 1 : 0;
    dy = (k == 2 || k == 4 || k == 6 || k == 7) ? 1 : 0;
    dz = (k == 3 || k == 5 || k == 6 || k == 7) ? 1 : 0;
}

// Index modulo N in 0..N-1 for every int (the reference's `%` for the
// positions inside the box, which give 0..N; positions outside it, which the
// reference indexes out of bounds, wrap here instead).
__device__ __forceinline__ size_t wrap_index(int i, int N)
{
    const int r = i % N;
    return static_cast<size_t>(r < 0 ? r + N : r);
}

__device__ __forceinline__ size_t node(const Cell &c, int dx, int dy, int dz, int N)
{
    const size_t a = wrap_index(c.i[0] + dx, N), b = wrap_index(c.i[1] + dy, N),
                 e = wrap_index(c.i[2] + dz, N);
    return (a * N + b) * N + e;
}

// Sort key of a particle: its base cell, wrapped into the grid (the cell
// whose corner (0,0,0) the particle deposits to).
__global__ void k_pm_cell_keys(const double *__restrict__ x, int n, int N, double c0, double c1,
                               double c2, double bw, double inv, long long *__restrict__ key,
                               int *__restrict__ idx)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const Cell c = cell_of(&x[3 * static_cast<size_t>(p)], c0, c1, c2, bw, inv);
    key[p] = static_cast<long long>(node(c, 0, 0, 0, N));
    idx[p] = p;
}

__global__ void k_pm_clear_ranges(int2 *__restrict__ r, size_t count)
{
    const size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (q < count) r[q] = make_int2(0, 0);
}

// [first, end) of each occupied cell's run in the sorted keys (other cells
// keep the empty range [0, 0)).
__global__ void k_pm_cell_ranges(const long long *__restrict__ key, int n, int2 *__restrict__ range)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const long long k = key[p];
    if (p == 0 || key[p - 1] != k) range[k].x = p;
    if (p == n - 1 || key[p + 1] != k) range[k].y = p + 1;
}

// Weights and mass of each particle in sorted order: (wx, wy, wz, m), the
// fractional offsets of cell_of (the weight of the far corner on each axis).
__global__ void k_pm_sorted_weights(const double *__restrict__ x, const double *__restrict__ m,
                                    const int *__restrict__ sidx, int n, double c0, double c1,
                                    double c2, double bw, double inv, double4 *__restrict__ wm)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int p = sidx[t];
    const Cell c = cell_of(&x[3 * static_cast<size_t>(p)], c0, c1, c2, bw, inv);
    wm[t] = make_double4(c.w[0], c.w[1], c.w[2], m[p]);
}

// Cloud in cell (src/acceleration_PM.c:15-105) as a gather: the mass at
// node q is the sum over the 8 corners d, in the reference's order, of the
// particles whose base cell is q - d, each adding m * wx * wy * wz for that
// corner (w or 1 - w per axis) -- the same product as the reference's for
// that particle and node.  A thread sums the nodes (a, b, 2j) and
// (a, b, 2j + 1), whose corners share 4 of their base cells: 12 range
// loads, all issued before any particle is read, and each node's sum in the
// same order as alone.
__global__ void k_cic_gather(const double4 *__restrict__ wm, const int2 *__restrict__ range, int N,
                             double *__restrict__ grid)
{
    const unsigned half = (static_cast<unsigned>(N) + 1) / 2;
    const unsigned q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= static_cast<unsigned>(N) * N * half) return;
    const int a = static_cast<int>(q / (static_cast<unsigned>(N) * half));
    const int b = static_cast<int>((q / half) % N);
    const int e0 = 2 * static_cast<int>(q % half);
    const bool two = e0 + 1 < N;
    // r[dx + 2 dy][z]: base cells (a - dx, b - dy, e0 - 1 + z), z = 0, 1, 2
    int2 r[4][3];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const size_t row = (wrap_index(a - (u & 1), N) * N + wrap_index(b - (u >> 1), N)) * N;
        r[u][0] = range[row + wrap_index(e0 - 1, N)];
        r[u][1] = range[row + static_cast<size_t>(e0)];
        r[u][2] = two ? range[row + static_cast<size_t>(e0 + 1)] : make_int2(0, 0);
    }
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        int dx, dy, dz;
        corner(k, dx, dy, dz);
        const int u = dx + 2 * dy;
        const int2 r0 = r[u][1 - dz], r1 = r[u][2 - dz];  // bases of e0 - dz, e0 + 1 - dz
        for (int t = r0.x; t < r0.y; t++) {
            const double4 w = wm[t];
            s0 += w.w * (dx ? w.x : 1.0 - w.x) * (dy ? w.y : 1.0 - w.y) * (dz ? w.z : 1.0 - w.z);
        }
        for (int t = r1.x; t < r1.y; t++) {
            const double4 w = wm[t];
            s1 += w.w * (dx ? w.x : 1.0 - w.x) * (dy ? w.y : 1.0 - w.y) * (dz ? w.z : 1.0 - w.z);
        }
    }
    const size_t q0 = (static_cast<size_t>(a) * N + b) * N + e0;
    grid[q0] = s0;
    if (two) grid[q0 + 1] = s1;
}

// Green's function of the Poisson equation on the half spectrum
// (src/acceleration_PM.c:313-335), with the potential's scale and the
// density normalisation folded into `factor`.
__global__ void k_pm_green(double2 *__restrict__ spec, int N, double two_pi_l2, double factor)
{
    const int H = N / 2 + 1;
    const size_t count = static_cast<size_t>(N) * N * H;
    const size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (q >= count) return;
    const int k = static_cast<int>(q % H);
    const int j = static_cast<int>((q / H) % N);
    const int i = static_cast<int>(q / (static_cast<size_t>(H) * N));
    const int kx = i < H ? i : i - N, ky = j < H ? j : j - N, kz = k;
    const double ksq = static_cast<double>(kx * kx + ky * ky + kz * kz) * two_pi_l2;
    const double kernel = ksq == 0.0 ? 0.0 : -1.0 / ksq;
    double2 s = spec[q];
    s.x = (s.x * kernel) * factor;
    s.y = (s.y * kernel) * factor;
    spec[q] = s;
}

// +grad phi at a particle's 8 CIC nodes (src/acceleration_PM.c:212-272: the
// same 4th-order expression per node and axis) and the CIC gather of those
// values (:107-210), one particle per lane.  Along each axis the stencils of
// the two nodes share one line of 6 potential values.
__global__ void k_cic_grad_interp(const double *__restrict__ x, const int *__restrict__ sidx, int n,
                                  int N, double c0, double c1, double c2, double bw, double inv,
                                  const double *__restrict__ phi, double twelve_h,
                                  double *__restrict__ a)
{
    const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (t0 >= n) return;
    // particles in the deposit's cell order: neighbouring threads read
    // neighbouring stencil points whatever the input order
    const int p = sidx[t0];
    const Cell c = cell_of(&x[3 * static_cast<size_t>(p)], c0, c1, c2, bw, inv);
    const size_t N2 = static_cast<size_t>(N) * N;
    size_t li[6], lj[6], lk[6];  // node indices i-2 .. i+3 along each axis
#pragma unroll
    for (int t = 0; t < 6; t++) {
        li[t] = wrap_index(c.i[0] + t - 2, N) * N2;
        lj[t] = wrap_index(c.i[1] + t - 2, N) * static_cast<size_t>(N);
        lk[t] = wrap_index(c.i[2] + t - 2, N);
    }
    double g[8][3];  // node (dx, dy, dz) at g[dx + 2 dy + 4 dz]
#pragma unroll
    for (int u = 0; u < 2; u++)
#pragma unroll
        for (int v = 0; v < 2; v++) {
            double lx[6], ly[6], lz[6];
#pragma unroll
            for (int t = 0; t < 6; t++) {
                lx[t] = phi[li[t] + lj[2 + u] + lk[2 + v]];
                ly[t] = phi[li[2 + u] + lj[t] + lk[2 + v]];
                lz[t] = phi[li[2 + u] + lj[2 + v] + lk[t]];
            }
#pragma unroll
            for (int e = 0; e < 2; e++) {
                g[e + 2 * u + 4 * v][0] = (lx[e] - 8.0 * lx[e + 1] + 8.0 * lx[e + 3] - lx[e + 4]) / twelve_h;
                g[u + 2 * e + 4 * v][1] = (ly[e] - 8.0 * ly[e + 1] + 8.0 * ly[e + 3] - ly[e + 4]) / twelve_h;
                g[u + 2 * v + 4 * e][2] = (lz[e] - 8.0 * lz[e + 1] + 8.0 * lz[e + 3] - lz[e + 4]) / twelve_h;
            }
        }
    double s[3];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        int dx, dy, dz;
        corner(k, dx, dy, dz);
        const double *gk = g[dx + 2 * dy + 4 * dz];
        const double wx = dx ? c.w[0] : c.wm[0], wy = dy ? c.w[1] : c.wm[1], wz = dz ? c.w[2] : c.wm[2];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double t = gk[d] * wx * wy * wz;
            s[d] = k == 0 ? t : s[d] + t;
        }
    }
    a[3 * static_cast<size_t>(p)] = s[0];
    a[3 * static_cast<size_t>(p) + 1] = s[1];
    a[3 * static_cast<size_t>(p) + 2] = s[2];
}

// Leapfrog updates of src/integrator_cosmology_leapfrog.c:100-196.
__global__ void k_cos_scale(double *__restrict__ out, const double *__restrict__ in, double s,
                            size_t n3)
{
    const size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (q < n3) out[q] = s * in[q];
}

__global__ void k_cos_kick(double *__restrict__ p, const double *__restrict__ a, double half_dt,
                           double da, size_t n3)
{
    const size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (q < n3) p[q] -= half_dt * a[q] / da;
}

// One pass per step of the cosmological Leapfrog
// (src/integrator_cosmology_leapfrog.c:150-200): the previous step's closing
// kick (when `pend`), this step's opening kick, the drift and the periodic
// wrap, each with the same operations as the separate passes.
__global__ void k_cos_kick_drift_wrap(double *__restrict__ p, double *__restrict__ x,
                                      const double *__restrict__ a, int pend, double half_prev,
                                      double da_prev, double half_dt, double da, double dt,
                                      double denom, size_t n3, double c0, double c1, double c2,
                                      double w)
{
    const size_t q = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (q >= n3) return;
    const double aq = a[q];
    double pq = p[q];
    if (pend) pq -= half_prev * aq / da_prev;
    pq -= half_dt * aq / da;
    p[q] = pq;
    const int d = static_cast<int>(q % 3);
    x[q] = wrap_coord(x[q] + dt * pq / denom, d == 0 ? c0 : (d == 1 ? c1 : c2), w);
}

ErrorStatus fft_error(hipfftResult r, const char *what)
{
    return WRAP_RAISE_ERROR_FMT(r == HIPFFT_ALLOC_FAILED ? GRAV_MEMORY_ERROR : GRAV_FAILURE,
                                "hipFFT call failed: %s (%d)", what, static_cast<int>(r));
}

#define GS_FFT(call)                                                                \
    do {                                                                            \
        hipfftResult gs_r_ = (call);                                                \
        if (gs_r_ != HIPFFT_SUCCESS) return fft_error(gs_r_, #call);                \
    } while (0)

}  // namespace

// Grid buffers and transform plans of one grid size.
struct PmWork {
    Context *c = nullptr;
    int N = 0;
    double *grid = nullptr;   // deposited density, then the potential (N^3)
    double2 *spec = nullptr;  // half spectrum N x N x (N/2 + 1)
    int2 *range = nullptr;    // [first, end) of each cell's particles in sorted order (N^3)
    int passes = 0;           // radix passes for keys below N^3
    // (cell, particle) sort buffers for up to `cap` particles
    int cap = 0;
    long long *key[2] = {nullptr, nullptr};
    int *idx[2] = {nullptr, nullptr};
    double4 *wm = nullptr;    // (wx, wy, wz, m) in sorted order
    unsigned *hist = nullptr;
    hipfftHandle r2c = 0, c2r = 0;
    bool have_r2c = false, have_c2r = false;

    void free_sort()
    {
        for (int k = 0; k < 2; k++) {
            if (key[k]) (void)hipFree(key[k]);
            if (idx[k]) (void)hipFree(idx[k]);
            key[k] = nullptr;
            idx[k] = nullptr;
        }
        if (hist) (void)hipFree(hist);
        if (wm) (void)hipFree(wm);
        hist = nullptr;
        wm = nullptr;
        cap = 0;
    }

    ErrorStatus ensure_particles(int n)
    {
        if (n <= cap) return make_success_error_status();
        GS_HIP(hipStreamSynchronize(c->stream));
        free_sort();
        const size_t hn = radix_hist_elements(n);
        for (int k = 0; k < 2; k++) {
            GS_HIP(hipMalloc(reinterpret_cast<void **>(&key[k]), sizeof(long long) * n));
            GS_HIP(hipMalloc(reinterpret_cast<void **>(&idx[k]), sizeof(int) * n));
        }
        GS_HIP(hipMalloc(reinterpret_cast<void **>(&wm), sizeof(double4) * n));
        GS_HIP(hipMalloc(reinterpret_cast<void **>(&hist), sizeof(unsigned) * hn));
        cap = n;
        return make_success_error_status();
    }

    ~PmWork()
    {
        if (c) (void)hipStreamSynchronize(c->stream);
        if (have_r2c) (void)hipfftDestroy(r2c);
        if (have_c2r) (void)hipfftDestroy(c2r);
        if (grid) (void)hipFree(grid);
        if (spec) (void)hipFree(spec);
        if (range) (void)hipFree(range);
        free_sort();
    }

    ErrorStatus init(Context *ctx, int n)
    {
        c = ctx;
        N = n;
        const size_t n3 = static_cast<size_t>(n) * n * n;
        const size_t nh = static_cast<size_t>(n) * n * (n / 2 + 1);
        GS_HIP(hipMalloc(reinterpret_cast<void **>(&grid), sizeof(double) * n3));
        GS_HIP(hipMalloc(reinterpret_cast<void **>(&spec), sizeof(double2) * nh));
        GS_HIP(hipMalloc(reinterpret_cast<void **>(&range), sizeof(int2) * n3));
        int bits = 0;
        while ((1ull << bits) < n3) bits++;
        passes = bits > 8 ? (bits + 7) / 8 : 1;
        GS_FFT(hipfftPlan3d(&r2c, n, n, n, HIPFFT_D2Z));
        have_r2c = true;
        GS_FFT(hipfftPlan3d(&c2r, n, n, n, HIPFFT_Z2D));
        have_c2r = true;
        GS_FFT(hipfftSetStream(r2c, c->stream));
        GS_FFT(hipfftSetStream(c2r, c->stream));
        return make_success_error_status();
    }
};

namespace {

// The buffers and plans of the last grid size used, kept for the next call
// (plans take milliseconds to build; a 256^3 grid holds 270 MB).  Used under
// the context mutex.
PmWork *g_pm = nullptr;

ErrorStatus pm_work(Context *c, int N, PmWork **out)
{
    if (!g_pm || g_pm->N != N) {
        delete g_pm;
        g_pm = new PmWork;
        ErrorStatus st = g_pm->init(c, N);
        if (st.return_code != GRAV_SUCCESS) {
            delete g_pm;
            g_pm = nullptr;
            return st;
        }
    }
    *out = g_pm;
    return make_success_error_status();
}

}  // namespace

// The PM acceleration of n particles (device arrays) into a (device).
ErrorStatus pm_acceleration(PmWork &W, const double *x, const double *m, int n,
                            const double *center, double box_width, double H0, double omega_m,
                            double mean_bkg_density, double scale_factor, double *a)
{
    Context *c = W.c;
    const int N = W.N;
    const size_t n3g = static_cast<size_t>(N) * N * N;
    const size_t nh = static_cast<size_t>(N) * N * (N / 2 + 1);
    const double box_length = box_width * 2.0;
    const double bw = box_length / 2.0;
    const double cell = box_length / N;
    const double inv = 1.0 / cell;
    if (n <= 0) return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "Number of particles must be positive. Got: %d", n);
    GS_TRY(W.ensure_particles(n));
    int rec = timing_begin(c, "pm_deposit");
    hipLaunchKernelGGL(k_pm_cell_keys, grid_for(n), dim3(PM_BLOCK), 0, c->stream, x, n, N, center[0],
                       center[1], center[2], bw, inv, W.key[0], W.idx[0]);
    GS_HIP(hipGetLastError());
    long long *skey;
    int *sidx;
    GS_TRY(radix_sort_pairs(c, W.key[0], W.idx[0], W.key[1], W.idx[1], n, W.passes, W.hist,
                            &skey, &sidx));
    hipLaunchKernelGGL(k_pm_clear_ranges, grid_for(n3g), dim3(PM_BLOCK), 0, c->stream, W.range, n3g);
    hipLaunchKernelGGL(k_pm_cell_ranges, grid_for(n), dim3(PM_BLOCK), 0, c->stream, skey, n, W.range);
    hipLaunchKernelGGL(k_pm_sorted_weights, grid_for(n), dim3(PM_BLOCK), 0, c->stream, x, m, sidx, n,
                       center[0], center[1], center[2], bw, inv, W.wm);
    hipLaunchKernelGGL(k_cic_gather, grid_for(static_cast<size_t>(N) * N * ((N + 1) / 2)), dim3(PM_BLOCK), 0,
                       c->stream, W.wm, W.range, N, W.grid);
    timing_end(c, rec);
    GS_HIP(hipGetLastError());
    rec = timing_begin(c, "pm_fft");
    GS_FFT(hipfftExecD2Z(W.r2c, W.grid, reinterpret_cast<hipfftDoubleComplex *>(W.spec)));
    const double two_pi_l2 = 4.0 * M_PI * M_PI / (box_length * box_length);
    const double factor = ((3.0 / 2.0) * omega_m * H0 * H0) / (scale_factor * static_cast<double>(n3g));
    const double cell_volume = cell * cell * cell;
    hipLaunchKernelGGL(k_pm_green, grid_for(nh), dim3(PM_BLOCK), 0, c->stream, W.spec, N,
                       two_pi_l2, factor / (cell_volume * mean_bkg_density));
    GS_HIP(hipGetLastError());
    GS_FFT(hipfftExecZ2D(W.c2r, reinterpret_cast<hipfftDoubleComplex *>(W.spec), W.grid));
    timing_end(c, rec);
    rec = timing_begin(c, "pm_interp");
    hipLaunchKernelGGL(k_cic_grad_interp, grid_for(n), dim3(PM_BLOCK), 0, c->stream, x, sidx, n, N,
                       center[0], center[1], center[2], bw, inv, W.grid, 12.0 * cell, a);
    timing_end(c, rec);
    GS_HIP(hipGetLastError());
    return make_success_error_status();
}

ErrorStatus pm_wrap(Context *c, double *x, int n, const double *center, double box_width)
{
    const size_t n3 = static_cast<size_t>(n) * 3;
    hipLaunchKernelGGL(k_pm_wrap, grid_for(n3), dim3(PM_BLOCK), 0, c->stream, x, n3, center[0],
                       center[1], center[2], box_width);
    GS_HIP(hipGetLastError());
    return make_success_error_status();
}

namespace {

// 1024^3 grids need 43 GB of buffers; beyond that the reference's int
// indexing overflows anyway.
constexpr int PM_MAX_GRID = 1024;

ErrorStatus check_grid(int pm_grid_size)
{
    if (pm_grid_size <= 0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "pm_grid_size must be positive. Got: %d",
                                    pm_grid_size);
    if (pm_grid_size > PM_MAX_GRID)
        return WRAP_RAISE_ERROR_FMT(GRAV_NOT_IMPLEMENTED_ERROR,
                                    "pm_grid_size %d exceeds the supported %d", pm_grid_size,
                                    PM_MAX_GRID);
    return make_success_error_status();
}

}  // namespace

}  // namespace gs

using namespace gs;

extern "C" {

// src/cosmology.c:5-12
double compute_omega_k(const double omega_m, const double omega_lambda)
{
    return 1.0 - omega_m - omega_lambda;
}

// src/cosmology.c:14-26
double compute_da(const double a, const double H0, const double omega_m, const double omega_lambda)
{
    const double omega_k = compute_omega_k(omega_m, omega_lambda);
    return H0 * a * sqrt(omega_lambda + omega_m / (a * a * a) + omega_k / (a * a));
}

// src/system.c:121-141
CosmologicalSystem get_new_cosmological_system(void)
{
    CosmologicalSystem s;
    s.num_particles = 0;
    s.particle_ids = nullptr;
    s.x = nullptr;
    s.v = nullptr;
    s.m = nullptr;
    s.h = -1.0;
    s.omega_m = -1.0;
    s.omega_lambda = -1.0;
    s.omega_k = 0.0;
    s.box_center[0] = s.box_center[1] = s.box_center[2] = 0.0;
    s.box_width = -1.0;
    s.unit_mass = 1.0;
    s.unit_length = 1.0;
    s.unit_time = 1.0;
    return s;
}

void free_cosmological_system(CosmologicalSystem *s)
{
    free(s->particle_ids);
    free(s->x);
    free(s->v);
    free(s->m);
}

// src/system.c:143-172
ErrorStatus get_initialized_cosmological_system(CosmologicalSystem *s, const int num_particles)
{
    if (!s) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "System is NULL");
    *s = get_new_cosmological_system();
    s->num_particles = num_particles;
    const size_t n = num_particles > 0 ? static_cast<size_t>(num_particles) : 0;
    s->particle_ids = static_cast<int *>(malloc(n * sizeof(int)));
    s->x = static_cast<double *>(calloc(n * 3, sizeof(double)));
    s->v = static_cast<double *>(calloc(n * 3, sizeof(double)));
    s->m = static_cast<double *>(calloc(n, sizeof(double)));
    if (!s->particle_ids || !s->x || !s->v || !s->m) {
        free_cosmological_system(s);
        return WRAP_RAISE_ERROR(GRAV_MEMORY_ERROR, "Failed to allocate memory for system");
    }
    for (int i = 0; i < num_particles; i++) s->particle_ids[i] = i;
    return make_success_error_status();
}

// src/system.c:174-242
ErrorStatus finalize_cosmological_system(CosmologicalSystem *s)
{
    if (!s) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "System is NULL");
    if (s->num_particles <= 0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "Number of particles must be positive. Got: %d",
                                    s->num_particles);
    if (!s->particle_ids)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "System array particle_ids is NULL");
    if (!s->x) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "System array x is NULL");
    if (!s->v) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "System array v is NULL");
    if (!s->m) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "System array m is NULL");
    if (s->h <= 0.0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR,
                                    "Hubble parameter system->h must be positive. Got: %g", s->h);
    if (s->omega_m <= 0.0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "omega_m must be positive. Got: %g",
                                    s->omega_m);
    if (s->omega_lambda <= 0.0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "omega_lambda must be positive. Got: %g",
                                    s->omega_lambda);
    if (s->box_width <= 0.0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "Box width must be positive. Got: %g",
                                    s->box_width);
    return make_success_error_status();
}

// src/system.c:1337-1380, on the host arrays (the same formula as k_pm_wrap)
void set_periodic_boundary_conditions(CosmologicalSystem *s)
{
    const double w = s->box_width;
    for (int i = 0; i < s->num_particles; i++)
        for (int d = 0; d < 3; d++) {
            double &x = s->x[3 * static_cast<size_t>(i) + d];
            const double c = s->box_center[d];
            const double p = x - c;
            if (p < -w) {
                const double f = fmod(p, w);
                x = ((static_cast<int>(p / w) % 2 != 0) ? f + w : f) + c;
            } else if (p > w) {
                const double f = fmod(p, w);
                x = ((static_cast<int>(p / w) % 2 != 0) ? f - w : f) + c;
            }
        }
}

// src/acceleration_PM.c:274-391 on host arrays: x, m copied in, a copied out.
ErrorStatus acceleration_PM(double *a, const int num_particles, const double *x, const double *m,
                            const double *box_center, const double box_width, const double H0,
                            const double omega_m, const double mean_bkg_density,
                            const int pm_grid_size, const double scale_factor)
{
    if (!a || !x || !m || !box_center) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    if (num_particles <= 0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR,
                                    "Number of particles must be positive. Got: %d", num_particles);
    GS_TRY(check_grid(pm_grid_size));
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    const size_t n3 = static_cast<size_t>(num_particles) * 3;
    void *dx, *dm, *da;
    GS_TRY(slot(c, SLOT_X, sizeof(double) * n3, &dx));
    GS_TRY(slot(c, SLOT_M, sizeof(double) * num_particles, &dm));
    GS_TRY(slot(c, SLOT_A, sizeof(double) * n3, &da));
    GS_HIP(hipMemcpyAsync(dx, x, sizeof(double) * n3, hipMemcpyHostToDevice, c->stream));
    GS_HIP(hipMemcpyAsync(dm, m, sizeof(double) * num_particles, hipMemcpyHostToDevice, c->stream));
    PmWork *W;
    GS_TRY(pm_work(c, pm_grid_size, &W));
    GS_TRY(pm_acceleration(*W, static_cast<double *>(dx), static_cast<double *>(dm), num_particles,
                           box_center, box_width, H0, omega_m, mean_bkg_density, scale_factor,
                           static_cast<double *>(da)));
    GS_HIP(hipMemcpyAsync(a, da, sizeof(double) * n3, hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    GS_TRY(timing_flush(c));
    return make_success_error_status();
}

// src/output.c:295-349: cosmological snapshots are HDF5 only, which is not
// built in (as in a reference build without HDF5).
ErrorStatus output_snapshot_cosmology(OutputParam *op, const CosmologicalSystem *system,
                                      const SimulationStatus *status, const Settings *settings)
{
    (void)system;
    (void)status;
    (void)settings;
    if (!op) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    switch (op->method) {
        case OUTPUT_METHOD_DISABLED:
            break;
        case OUTPUT_METHOD_CSV:
            return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR,
                                    "CSV output method is not supported for cosmological simulation.");
        case OUTPUT_METHOD_HDF5:
            return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR,
                                    "HDF5 output method is not available. Please recompile with "
                                    "HDF5 support.");
        default:
            return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "Unknown output method. Got: %d",
                                        op->method);
    }
    op->output_count_++;
    return make_success_error_status();
}

// src/integrator_cosmology_leapfrog.c:27-255, device-resident: positions,
// momenta p = a^2 v and accelerations stay in HBM; the expansion rate and
// the step arithmetic run on the host as in the reference.  As there, the
// velocities are converted back from the momenta only for snapshots, so a
// run without output returns the caller's v unchanged.
ErrorStatus leapfrog_cosmology(CosmologicalSystem *system, OutputParam *op,
                               SimulationStatus *status, Settings *settings, double dt,
                               const double a_begin, const double a_final, const int pm_grid_size)
{
    if (!system || !op || !status || !settings)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    if (system->num_particles <= 0)
        return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR, "Number of particles must be positive");
    GS_TRY(check_grid(pm_grid_size));
    if (op->method != OUTPUT_METHOD_DISABLED)
        GS_TRY(output_snapshot_cosmology(op, system, status, settings));
    const int n = system->num_particles;
    const size_t n3 = static_cast<size_t>(n) * 3;
    const double H0 = system->h * 100.0;
    const double om = system->omega_m, ol = system->omega_lambda;
    const double L = system->box_width * 2;
    double msum = 0.0;  // vec_sum (src/math_functions.c:13-21)
    for (int i = 0; i < n; i++) msum += system->m[i];
    const double mean_bkg_density = msum / (L * L * L);

    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    double *mem = nullptr;
    GS_HIP(hipMalloc(reinterpret_cast<void **>(&mem), sizeof(double) * (4 * n3 + n)));
    struct Free {
        Context *c;
        double *p;
        ~Free()
        {
            (void)hipStreamSynchronize(c->stream);
            (void)hipFree(p);
        }
    } guard{c, mem};
    double *x = mem, *v = x + n3, *p = v + n3, *a = p + n3, *m = a + n3;
    GS_HIP(hipMemcpyAsync(x, system->x, sizeof(double) * n3, hipMemcpyHostToDevice, c->stream));
    GS_HIP(hipMemcpyAsync(v, system->v, sizeof(double) * n3, hipMemcpyHostToDevice, c->stream));
    GS_HIP(hipMemcpyAsync(m, system->m, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    PmWork *W;
    GS_TRY(pm_work(c, pm_grid_size, &W));
    const dim3 g3 = grid_for(n3), blk(PM_BLOCK);
    const double *ctr = system->box_center;

    double t = a_begin;
    status->t = t;
    GS_TRY(pm_wrap(c, x, n, ctr, system->box_width));
    hipLaunchKernelGGL(k_cos_scale, g3, blk, 0, c->stream, p, v, t * t, n3);
    GS_HIP(hipGetLastError());
    GS_TRY(pm_acceleration(*W, x, m, n, ctr, system->box_width, H0, om, mean_bkg_density, t, a));
    // (a step size of 0 would make the reference's step count infinite)
    const long long total =
        dt > 0.0 ? static_cast<long long>(ceil((a_final - a_begin) / dt)) : 0;
    status->dt = dt;
    status->num_steps = 0;
    // a step's closing kick is applied in the next step's first pass (or
    // after the loop)
    int pend = 0;
    double half_prev = 0.0, da_prev = 1.0;
    while (status->num_steps < total) {
        if (t + dt > a_final) dt = a_final - t;
        status->dt = dt;
        const double da0 = compute_da(t, H0, om, ol);
        t += 0.5 * dt;
        const double da = compute_da(t, H0, om, ol);
        hipLaunchKernelGGL(k_cos_kick_drift_wrap, g3, blk, 0, c->stream, p, x, a, pend, half_prev,
                           da_prev, 0.5 * dt, da0, dt, (t * t) * da, n3, ctr[0], ctr[1], ctr[2],
                           system->box_width);
        GS_HIP(hipGetLastError());
        GS_TRY(pm_acceleration(*W, x, m, n, ctr, system->box_width, H0, om, mean_bkg_density, t, a));
        pend = 1;
        half_prev = 0.5 * dt;
        da_prev = da;
        status->num_steps++;
        t = a_begin + status->num_steps * dt;
        status->t = t;
        if (settings->is_exit_ptr && *settings->is_exit_ptr) break;
        // keep the host a bounded number of steps ahead of the device
        if ((status->num_steps & 15) == 0) GS_HIP(hipStreamSynchronize(c->stream));
    }
    if (pend) {
        hipLaunchKernelGGL(k_cos_kick, g3, blk, 0, c->stream, p, a, half_prev, da_prev, n3);
        GS_HIP(hipGetLastError());
    }
    GS_HIP(hipMemcpyAsync(system->x, x, sizeof(double) * n3, hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    GS_TRY(timing_flush(c));
    return make_success_error_status();
}

// src/grav_sim.c:140-260 (a build with FFTW: the GPU transforms take its place)
ErrorStatus launch_cosmological_simulation(CosmologicalSystem *system, OutputParam *op,
                                           SimulationStatus *status, Settings *settings,
                                           double dt, const double a_begin, const double a_final,
                                           const int pm_grid_size)
{
    GS_TRY(finalize_cosmological_system(system));
    if (!op || !status || !settings) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    if (op->method == OUTPUT_METHOD_CSV)
        return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR,
                                "CSV output is not supported for cosmological simulations. Please "
                                "use HDF5 output format.");
    GS_TRY(finalize_output_param(op, settings));
    if (dt < 0.0) return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "dt must be non=negative. Got: %g", dt);
    if (a_begin <= 0.0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "a_begin must be positive. Got: %g", a_begin);
    if (a_final <= 0.0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "a_final must be positive. Got: %g", a_final);
    if (a_final < a_begin)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR,
                                    "a_final must be greater or equal to a_begin. Got: %g, %g",
                                    a_begin, a_final);
    if (pm_grid_size <= 0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "pm_grid_size must be positive. Got: %d",
                                    pm_grid_size);
    if (settings->verbose >= GRAV_VERBOSITY_IGNORE_INFO) fputs("Launching simulation...\n", stdout);
    return WRAP_TRACEBACK(
        leapfrog_cosmology(system, op, status, settings, dt, a_begin, a_final, pm_grid_size));
}

}  // extern "C"
