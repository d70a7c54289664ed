// whfast.hip -- WHFast on the GPU for ensembles of independent systems
// (src/integrator_whfast.c:200-1279, src/system.c:444-528,1210-1335).
//
// One workgroup integrates one system for a chunk of steps; the system's
// arrays stay in device memory between chunks.  Every FP64 expression keeps
// the reference's association order (contraction is off for this file), and
// every sum runs in the reference's index order, so a system's trajectory is
// bitwise equal to the reference's (with the particle-0 acceleration slot the
// reference never writes taken as zero).
//
// Per step:  distances |jx_i - jx_primary| -> sort of (distance, index) in
// LDS from the previous order (odd-even passes, bitonic fallback) -> move
// the particles that changed place -> eta (cumulative mass over the
// massive particles; massless ones add exactly 0) -> universal-variable
// Kepler drift per particle -> removal of unconverged particles (ascending
// order) -> Jacobi->Cartesian positions (the recurrence walked over the
// massive particles and the heads of the massless runs, then filled in
// parallel) -> massless-form accelerations and kick.  Cartesian
// velocities are only needed at the end and are reconstructed there.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "gs_internal.h"

#pragma clang fp contract(off)

namespace gs {

constexpr int WF_BLOCK = 256;
constexpr int WF_MAX_N = 2048;      // particles per system (LDS sort capacity)
constexpr int WF_MAX_MASSIVE = 64;  // massive particles per system
constexpr int WF_PAIR_NM = 16;      // massive particles up to which pair terms are tabulated
constexpr int WF_PAIRS = WF_PAIR_NM * (WF_PAIR_NM - 1) / 2;

struct WfSys {
    long long off;  // first particle in the flat arrays
    int cap;        // particles at the start
};

struct WfArgs {
    const WfSys *sys;
    double *jx, *jv, *x, *v, *m, *jx2, *jv2, *m2;
    int *ids, *ids2;
    int *count;       // live particles per system
    int *status;      // per system: bit 0 = Stumpff argument not finite (no removal)
    double G, eps, dt, dt_last;
    long long step0, nsteps, total_steps;
    int snapshot_last;  // the launch's last step is a snapshot step
    int method;       // 1 pairwise form, 2 massless form
    int remove_invalid;
};

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ double norm3(double a, double b, double c)
{
    return __builtin_sqrt(a * a + b * b + c * c);
}

// z / b for the constant divisors of the Stumpff series, correctly rounded
// with one multiply and two FMAs instead of a full division (Markstein's
// theorem: with y = RN(1/b) and q = RN(z y) within an ulp of z / b, the
// remainder r = z - b q is exact and RN(q + r y) = RN(z / b)).  Quotients
// near the subnormal range, where the remainder can be inexact, take the
// division itself.
__device__ __forceinline__ double div_const(double z, double b, double y)
{
    if (!(fabs(z) >= 0x1p-900)) return z / b;
    const double q = z * y;
    const double r = __builtin_fma(-q, b, z);
    return __builtin_fma(r, y, q);
}

#define DIVC(z, b) div_const((z), (b), 1.0 / (b))

// src/integrator_whfast.c:774-815
__device__ void stumpff(double z, double &c0, double &c1, double &c2, double &c3)
{
    int k = 0;
    while (fabs(z) > 0.1) {
        z /= 4.0;
        k++;
    }
    // the reference's nested series, innermost factor first
    double t3 = 1.0 - DIVC(z, 210.0);
    t3 = 1.0 - DIVC(z, 156.0) * t3;
    t3 = 1.0 - DIVC(z, 110.0) * t3;
    t3 = 1.0 - DIVC(z, 72.0) * t3;
    t3 = 1.0 - DIVC(z, 42.0) * t3;
    t3 = 1.0 - DIVC(z, 20.0) * t3;
    t3 = DIVC(t3, 6.0);
    double t2 = 1.0 - DIVC(z, 182.0);
    t2 = 1.0 - DIVC(z, 132.0) * t2;
    t2 = 1.0 - DIVC(z, 90.0) * t2;
    t2 = 1.0 - DIVC(z, 56.0) * t2;
    t2 = 1.0 - DIVC(z, 30.0) * t2;
    t2 = 1.0 - DIVC(z, 12.0) * t2;
    t2 = t2 / 2.0;
    double t1 = 1.0 - z * t3;
    double t0 = 1.0 - z * t2;
    while (k > 0) {
        t3 = (t2 + t0 * t3) / 4.0;
        t2 = (t1 * t1) / 2.0;
        t1 = t0 * t1;
        t0 = (2.0 * t0 * t0) - 1.0;
        k--;
    }
    c0 = t0;
    c1 = t1;
    c2 = t2;
    c3 = t3;
}

// Kepler drift of one Jacobi particle (src/integrator_whfast.c:459-578).
// Returns 1 when the particle is to be removed (non-converged with error
// above 1e-5 or a non-finite Stumpff argument), 2 when the Stumpff argument
// was not finite (reported when removal is off).
__device__ int kepler_drift(double *jx, double *jv, double gm, double dt)
{
    const double x0 = jx[0], x1 = jx[1], x2 = jx[2];
    const double v0 = jv[0], v1 = jv[1], v2 = jv[2];
    const double xn = norm3(x0, x1, x2);
    const double vn = norm3(v0, v1, v2);
    const double rv = (x0 * v0 + x1 * v1 + x2 * v2) / xn;
    const double alpha = 2.0 * gm / xn - (vn * vn);
    double s = dt / xn;
    double c0 = 0.0, c1 = 0.0, c2 = 0.0, c3 = 0.0;
    bool converged = false, zbad = false;
    for (int it = 0; it < 500; it++) {
        const double z = alpha * (s * s);
        if (!isfinite(z)) {
            zbad = true;
            break;
        }
        stumpff(z, c0, c1, c2, c3);
        const double F = (xn * s * c1 + xn * rv * (s * s) * c2 + gm * (s * s * s) * c3 - dt);
        const double dF = (xn * c0 + xn * rv * s * c1 + gm * (s * s) * c2);
        const double ds = -F / dF;
        s += ds;
        if (fabs(ds) < 1e-12) {
            converged = true;
            break;
        }
    }
    const double r = xn * c0 + xn * rv * s * c1 + gm * (s * s) * c2;
    int flag = 0;
    if (!converged) {
        const double err = (xn * s * c1 + xn * rv * (s * s) * c2 + gm * (s * s * s) * c3 - dt) / r;
        if (err > 1e-5 || zbad) flag = 1;
        if (zbad) flag |= 2;
    }
    const double f = 1.0 - gm * (s * s) * c2 / xn;
    const double g = dt - gm * (s * s * s) * c3;
    const double df = -gm * s * c1 / (r * xn);
    const double dg = 1.0 - gm * (s * s) * c2 / r;
    jx[0] = f * x0 + g * v0;
    jx[1] = f * x1 + g * v1;
    jx[2] = f * x2 + g * v2;
    jv[0] = df * x0 + dg * v0;
    jv[1] = df * x1 + dg * v1;
    jv[2] = df * x2 + dg * v2;
    return flag;
}

__device__ __forceinline__ double cube_plus(double a, double b, double c, double eps3)
{
    const double n = norm3(a, b, c);
    return (n * n * n) + eps3;
}

// Block-wide exclusive scan of small integers (WF_BLOCK threads).
__device__ int block_scan_excl(int v, int *total, int *buf)
{
    const int t = threadIdx.x;
    buf[t] = v;
    __syncthreads();
    for (int off = 1; off < WF_BLOCK; off <<= 1) {
        const int add = t >= off ? buf[t - off] : 0;
        __syncthreads();
        buf[t] += add;
        __syncthreads();
    }
    const int incl = buf[t];
    *total = buf[WF_BLOCK - 1];
    __syncthreads();
    return incl - v;
}

struct WfShared {
    double key[WF_MAX_N];
    short idx[WF_MAX_N];
    int scan_buf[WF_BLOCK];
    int mlist[WF_MAX_MASSIVE];   // ordinals of massive particles (sorted order)
    double mprefix[WF_MAX_MASSIVE];  // eta at each massive particle
    double mx[WF_MAX_MASSIVE][3];    // cartesian positions of massive particles
    double mm[WF_MAX_MASSIVE];
    // straddle sums: massless particles in gap g (g massive particles inside)
    // and massive particle p (pairs with one member inside, one outside)
    double strad_l[WF_MAX_MASSIVE + 1][3];
    double strad_m[WF_MAX_MASSIVE][3];
    double pterm[WF_PAIRS][3];  // G m_q m_r (x_r - x_q) / (|x_r - x_q|^3 + eps^3), q < r
    unsigned mbits[WF_MAX_N / 32];  // massive flags, one bit per particle
    int nm;
    int n;
    int any_remove;
    int bad;  // bit d: some jq component d is not finite
    int primary;
};
static_assert(WF_MAX_N == 64 * 32, "one 32-bit mask word per lane of one wavefront");

// LDS reused between the phases of a step: removal flags while drifting,
// then the Jacobi -> Cartesian quotients.
constexpr int J2C_K = 3;  // explicit quotients per massless run before its fixed point
union WfScratch {
    int flags[WF_MAX_N];
    struct {
        double exp[WF_MAX_MASSIVE][J2C_K][3];  // run below massive ordinal g: first quotients
        double mq[WF_MAX_MASSIVE][3];          // quotient at massive ordinal g
        double mjq[WF_MAX_MASSIVE][3];         // jq at massive ordinal g
        double q0[3];                          // Cartesian coordinate of particle 0
        unsigned char nexp[WF_MAX_MASSIVE][3];  // quotients stored for the run of gap g
        int fallback[3];                       // component d needs the full serial chain
    } j;
};
static_assert(sizeof(WfScratch) == sizeof(int) * WF_MAX_N, "the quotients fit in the flags");

// Sort (distance, index) ascending, lexicographically, for n <= WF_MAX_N.
__device__ void bitonic_sort(WfShared &S, int n)
{
    int P = 1;
    while (P < n) P <<= 1;
    for (int i = n + threadIdx.x; i < P; i += WF_BLOCK) {
        S.key[i] = INFINITY;
        S.idx[i] = static_cast<short>(i);
    }
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P; i += WF_BLOCK) {
                const int l = i ^ j;
                if (l > i) {
                    const double ki = S.key[i], kl = S.key[l];
                    const short ii = S.idx[i], il = S.idx[l];
                    const bool gt = (ki > kl) || (ki == kl && ii > il);
                    const bool up = (i & k) == 0;
                    if (gt == up) {
                        S.key[i] = kl;
                        S.key[l] = ki;
                        S.idx[i] = il;
                        S.idx[l] = ii;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Odd-even transposition passes tried before the bitonic sort.
constexpr int WF_OET_PASSES = 24;

// Sort the (distance, index) pairs in S.key / S.idx (idx = 0..n-1 on entry)
// ascending, lexicographically.  Each step arrives in the previous step's
// order, so the distances are nearly sorted: odd-even transposition passes
// finish in a few passes; if they do not within WF_OET_PASSES, the bitonic
// sort takes the partly sorted pairs.  The pairs are distinct, so both give
// the same order.  NaN distances go straight to the bitonic sort.
__device__ void sort_pairs(WfShared &S, int n)
{
    int unsorted = 0, nan = 0;
    for (int i = threadIdx.x; i < n; i += WF_BLOCK) {
        const double k = S.key[i];
        if (k != k) nan = 1;
        if (i + 1 < n && k > S.key[i + 1]) unsorted = 1;
    }
    if (__syncthreads_or(nan)) {
        bitonic_sort(S, n);
        return;
    }
    if (!__syncthreads_or(unsorted)) return;
    bool prev = true;
    for (int pass = 0; pass < WF_OET_PASSES; pass++) {
        int sw = 0;
        for (int i = (pass & 1) + 2 * static_cast<int>(threadIdx.x); i + 1 < n; i += 2 * WF_BLOCK) {
            const double ki = S.key[i], kl = S.key[i + 1];
            const short ii = S.idx[i], il = S.idx[i + 1];
            if (ki > kl || (ki == kl && ii > il)) {
                S.key[i] = kl;
                S.key[i + 1] = ki;
                S.idx[i] = il;
                S.idx[i + 1] = ii;
                sw = 1;
            }
        }
        const bool any = __syncthreads_or(sw);
        if (!any && !prev) return;  // both parities left every pair in order
        prev = any;
    }
    bitonic_sort(S, n);
}

// Distances from the particle whose id is 0 (src/system.c:1232-1277) and
// their sort: S.idx[i] = the current place of the particle that goes to
// sorted place i (the identity when there is no particle with id 0).
__device__ void sort_keys(WfShared &S, int n, const double *pos, const int *ids)
{
    if (threadIdx.x == 0) {
        int pr = -1;
        if (n > 0 && ids[0] == 0)
            pr = 0;
        else
            for (int i = 0; i < n; i++)
                if (ids[i] == 0) {
                    pr = i;
                    break;
                }
        S.primary = pr;
    }
    __syncthreads();
    const int pr = S.primary;
    if (pr < 0) {  // no particle with id 0: the reference reports an error; keep order
        for (int i = threadIdx.x; i < n; i += WF_BLOCK) S.idx[i] = static_cast<short>(i);
        __syncthreads();
        return;
    }
    const double p0 = pos[3 * pr], p1 = pos[3 * pr + 1], p2 = pos[3 * pr + 2];
    for (int i = threadIdx.x; i < n; i += WF_BLOCK) {
        S.key[i] = i == pr ? 0.0 : norm3(pos[3 * i] - p0, pos[3 * i + 1] - p1, pos[3 * i + 2] - p2);
        S.idx[i] = static_cast<short>(i);
    }
    __syncthreads();
    sort_pairs(S, n);
}

// sort_keys, then permute (pos, vel, m, ids) through the alternate buffers;
// only the particles that changed place move.
__device__ void sort_by_distance(WfShared &S, int n, double *pos, double *vel, double *m,
                                 int *ids, double *pos2, double *vel2, double *m2, int *ids2)
{
    sort_keys(S, n, pos, ids);
    for (int i = threadIdx.x; i < n; i += WF_BLOCK) {
        const int s = S.idx[i];
        if (s == i) continue;
        pos2[3 * i] = pos[3 * s];
        pos2[3 * i + 1] = pos[3 * s + 1];
        pos2[3 * i + 2] = pos[3 * s + 2];
        vel2[3 * i] = vel[3 * s];
        vel2[3 * i + 1] = vel[3 * s + 1];
        vel2[3 * i + 2] = vel[3 * s + 2];
        m2[i] = m[s];
        ids2[i] = ids[s];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += WF_BLOCK) {
        if (S.idx[i] == i) continue;
        pos[3 * i] = pos2[3 * i];
        pos[3 * i + 1] = pos2[3 * i + 1];
        pos[3 * i + 2] = pos2[3 * i + 2];
        vel[3 * i] = vel2[3 * i];
        vel[3 * i + 1] = vel2[3 * i + 1];
        vel[3 * i + 2] = vel2[3 * i + 2];
        m[i] = m2[i];
        ids[i] = ids2[i];
    }
    __syncthreads();
}

// Massive list (m != 0, or every particle for the pairwise form) and the
// cumulative masses eta at each massive particle, summed in index order
// (src/integrator_whfast.c:1266-1279; massless particles add exactly 0).
__device__ void massive_list(WfShared &S, int n, const double *m, int method,
                             const short *perm = nullptr)
{
    // the mass at sorted place i: m[i], or m[perm[i]] before the arrays are permuted
    auto mass = [&](int i) { return m[perm ? perm[i] : i]; };
    // ordinals of massive particles, ascending: bit mask, then one wavefront
    // turns the mask words into the list (lane w owns particles 32w..32w+31)
    for (int w = threadIdx.x; w < WF_MAX_N / 32; w += WF_BLOCK) S.mbits[w] = 0u;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += WF_BLOCK)
        // particle 0 (the primary) always anchors the list
        if (i == 0 || method == 1 || mass(i) != 0.0) atomicOr(&S.mbits[i >> 5], 1u << (i & 31));
    __syncthreads();
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        unsigned b = S.mbits[lane];
        const int cnt = __popc(b);
        int incl = cnt;
        for (int off = 1; off < 64; off <<= 1) {
            const int u = __shfl_up(incl, off, 64);
            if (lane >= off) incl += u;
        }
        int pos = incl - cnt;
        while (b) {
            const int bit = __ffs(b) - 1;
            b &= b - 1u;
            if (pos < WF_MAX_MASSIVE) S.mlist[pos] = 32 * lane + bit;
            pos++;
        }
        if (lane == 63) S.nm = incl;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nm = S.nm < WF_MAX_MASSIVE ? S.nm : WF_MAX_MASSIVE;
        double e = 0.0;
        for (int q = 0; q < nm; q++) {
            const int i = S.mlist[q];
            e = (i == 0) ? mass(0) : e + mass(i);
            S.mprefix[q] = e;
            S.mm[q] = mass(i);
        }
    }
    __syncthreads();
}

// The largest massive ordinal g with mlist[g] <= i for the ascending i of
// one thread: the ordinal and the next massive place stay in registers and
// advance only when i passes that place (mlist[0] = 0 starts it).
struct MassiveCursor {
    int g = 0, next = 0;
    __device__ __forceinline__ void init(const WfShared &S, int nm)
    {
        g = 0;
        next = nm > 1 ? S.mlist[1] : 0x7fffffff;
    }
    __device__ __forceinline__ int at(const WfShared &S, int nm, int i)
    {
        while (i >= next) {
            g++;
            next = g + 1 < nm ? S.mlist[g + 1] : 0x7fffffff;
        }
        return g;
    }
};

// eta[i] for every particle, into S.key (free between the sorts).
__device__ void fill_eta(WfShared &S, int n)
{
    const int nm = S.nm < WF_MAX_MASSIVE ? S.nm : WF_MAX_MASSIVE;
    MassiveCursor mc;
    mc.init(S, nm);
    for (int i = threadIdx.x; i < n; i += WF_BLOCK) S.key[i] = S.mprefix[mc.at(S, nm, i)];
    __syncthreads();
}

// Jacobi -> Cartesian for one component (src/integrator_whfast.c:726-772),
// eta from S.key.  For a massless particle the recurrence is
//   c <- eta * ((c - 0 * jq) / eta),   q = jq + (c - 0 * jq) / eta
// with an unchanged eta, so once c maps onto itself it stays there for the
// rest of the run of massless particles: the loop then only forms q from the
// cached quotient, which is bitwise what the full recurrence produces.
__device__ void j2c_component(int n, int d, const double *jq, double *q, const double *m,
                              const WfShared &S)
{
    const double *eta = S.key;
    double c = eta[n - 1] * jq[d];
    bool fixed = false;
    double cq = 0.0;
    for (int i = n - 1; i > 0; i--) {
        const double mi = m[i];
        const double jqi = jq[3 * i + d];
        if (fixed && mi == 0.0 && isfinite(jqi)) {
            q[3 * i + d] = jqi + cq;
            continue;
        }
        const double ei = eta[i];
        const double ei1 = eta[i - 1];
        const double cn = (c - mi * jqi) / ei;
        q[3 * i + d] = jqi + cn;
        const double c2 = ei1 * cn;
        fixed = (mi == 0.0) && (ei == ei1) && (c2 == c) && (c != 0.0);
        cq = cn;
        c = c2;
    }
    q[d] = c / m[0];
}

// Quotient of the recurrence for the massless particle i in the run above
// massive ordinal g (its first J2C_K quotients, then the fixed point).
__device__ __forceinline__ double run_quotient(const WfShared &S, const WfScratch &X, int nm, int n,
                                               int g, int i, int d)
{
    const int top = (g + 1 < nm ? S.mlist[g + 1] : n) - 1;
    const int r = top - i, k = X.j.nexp[g][d];
    return X.j.exp[g][r < k ? r : k - 1][d];
}

// Jacobi -> Cartesian for the whole workgroup, first half.  The recurrence
// above only needs jq at the massive particles: inside a run of massless
// particles (finite jq, c != 0) the quotient is c / eta and c becomes
// eta (c / eta), which reaches a fixed point within a few particles.  Lane d
// therefore walks the massive particles and the head of each run (at most
// J2C_K quotients per run) and stores the quotients, with the coordinate of
// particle 0; then q = jq + quotient for every other particle, in parallel
// (j2c_fill, or fused into the kick).  Components with bit d of `bad` set (a
// non-finite jq), or whose walk meets c == 0 or no fixed point within J2C_K,
// are marked in X.j.fallback for the serial recurrence; both give the same
// bits.  Called by all threads.
__device__ void j2c_quotients(int n, const double *jq, const WfShared &S, WfScratch &X, int bad)
{
    const int tid = threadIdx.x;
    const int nm = S.nm < WF_MAX_MASSIVE ? S.nm : WF_MAX_MASSIVE;
    for (int q = tid; q < nm; q += WF_BLOCK) {
        const int j = S.mlist[q];
        X.j.mjq[q][0] = jq[3 * j];
        X.j.mjq[q][1] = jq[3 * j + 1];
        X.j.mjq[q][2] = jq[3 * j + 2];
    }
    __syncthreads();
    if (tid < 3) {
        const int d = tid;
        const double *eta = S.key;
        bool fail = ((bad >> d) & 1) != 0;
        double c = eta[n - 1] * X.j.mjq[0][d];  // (ordinal 0 is particle 0)
        for (int g = nm - 1; g >= 0 && !fail; g--) {
            const int lo = S.mlist[g];
            const int hi = (g + 1 < nm ? S.mlist[g + 1] : n) - 1;
            int k = 0;
            for (int i = hi; i > lo; i--) {
                if (k == J2C_K || c == 0.0) {
                    fail = true;
                    break;
                }
                // (c - 0 * jq) / eta with jq finite and c != 0 is c / eta
                const double cn = c / eta[i];
                X.j.exp[g][k][d] = cn;
                k++;
                const double c2 = eta[i - 1] * cn;
                if (c2 == c) break;  // the rest of the run repeats this quotient
                c = c2;
            }
            X.j.nexp[g][d] = static_cast<unsigned char>(k);
            if (!fail && g >= 1) {
                const double cn = (c - S.mm[g] * X.j.mjq[g][d]) / eta[lo];
                X.j.mq[g][d] = cn;
                c = eta[lo - 1] * cn;
            }
        }
        X.j.fallback[d] = fail ? 1 : 0;
        if (!fail) X.j.q0[d] = c / S.mm[0];
    }
    __syncthreads();
}

// Second half: q for every particle into global memory; the serial
// recurrence for the components marked in X.j.fallback.
__device__ void j2c_fill(int n, const double *jq, double *q, const double *m, const WfShared &S,
                         const WfScratch &X)
{
    const int tid = threadIdx.x;
    const int nm = S.nm < WF_MAX_MASSIVE ? S.nm : WF_MAX_MASSIVE;
    const bool f0 = X.j.fallback[0], f1 = X.j.fallback[1], f2 = X.j.fallback[2];
    if (tid < 3) {
        if (X.j.fallback[tid])
            j2c_component(n, tid, jq, q, m, S);
        else
            q[tid] = X.j.q0[tid];
    }
    MassiveCursor mc;
    mc.init(S, nm);
    for (int i = 1 + tid; i < n; i += WF_BLOCK) {
        const int g = mc.at(S, nm, i);
        const bool event = S.mlist[g] == i;
        for (int d = 0; d < 3; d++) {
            if (d == 0 ? f0 : (d == 1 ? f1 : f2)) continue;
            const double cn = event ? X.j.mq[g][d] : run_quotient(S, X, nm, n, g, i, d);
            q[3 * i + d] = jq[3 * i + d] + cn;
        }
    }
    __syncthreads();
}

// Both halves, with the non-finite check over the massless particles.
__device__ void j2c_block(int n, const double *jq, double *q, const double *m, WfShared &S,
                          WfScratch &X)
{
    if (threadIdx.x == 0) S.bad = 0;
    __syncthreads();
    int bad = 0;
    for (int i = 1 + threadIdx.x; i < n; i += WF_BLOCK) {
        if (m[i] != 0.0) continue;  // (runs hold massless particles only)
        for (int d = 0; d < 3; d++)
            if (!isfinite(jq[3 * i + d])) bad |= 1 << d;
    }
    if (bad) atomicOr(&S.bad, bad);
    __syncthreads();
    j2c_quotients(n, jq, S, X, S.bad);
    j2c_fill(n, jq, q, m, S, X);
}

// Cartesian -> Jacobi (src/integrator_whfast.c:682-724), one lane per
// component; particle 0 gets the centre of mass.
__device__ void c2j_component(int n, int d, const double *q, double *jq, const double *m,
                              const WfShared &S)
{
    const double *eta = S.key;
    double c = m[0] * q[d];
    for (int i = 1; i < n; i++) {
        const double ei1 = eta[i - 1];
        jq[3 * i + d] = q[3 * i + d] - c / ei1;
        c = c * (1.0 + m[i] / ei1) + m[i] * jq[3 * i + d];
    }
    jq[d] = c / eta[n - 1];
}

// Interaction accelerations and kick for particle i >= 1 (massless form,
// src/integrator_whfast.c:1022-1258; the pairwise form is the same with
// every particle in the massive list).
// xi, x0: Cartesian positions of particle i and of particle 0; jxi: its
// Jacobi position; ginside: for a massless particle, the number of massive
// particles inside it.
__device__ __forceinline__ void accel_kick_v(const WfShared &S, int i, bool massive, int p,
                                             const double *xi, const double *x0,
                                             const double *jxi, int ginside, double m0, double G,
                                             double eps3, double *jv, double dt,
                                             double *jv_sync = nullptr)
{
    double a[3];
    const double tv0 = xi[0] - x0[0], tv1 = xi[1] - x0[1], tv2 = xi[2] - x0[2];
    const double c0 = cube_plus(tv0, tv1, tv2, eps3);
    const double cj = cube_plus(jxi[0], jxi[1], jxi[2], eps3);
    const double ei = S.key[i], ei1 = S.key[i - 1];
    const double tv[3] = {tv0, tv1, tv2};
    if (massive) {
        for (int d = 0; d < 3; d++) a[d] = G * m0 * ei / ei1 * (jxi[d] / cj - tv[d] / c0);
    } else {
        for (int d = 0; d < 3; d++) a[d] = G * m0 * (jxi[d] / cj - tv[d] / c0);
    }
    const int nm = S.nm;
    double aux0 = 0.0, aux1 = 0.0, aux2 = 0.0;
    // interior massive particles (ordinal 1.., position < i)
    for (int q = 1; q < nm; q++) {
        const int j = S.mlist[q];
        if (massive ? q >= p : j >= i) break;
        const double t0 = xi[0] - S.mx[q][0], t1 = xi[1] - S.mx[q][1], t2 = xi[2] - S.mx[q][2];
        const double cc = cube_plus(t0, t1, t2, eps3);
        aux0 += G * S.mm[q] * t0 / cc;
        aux1 += G * S.mm[q] * t1 / cc;
        aux2 += G * S.mm[q] * t2 / cc;
    }
    if (massive) {
        a[0] -= aux0 * ei / ei1;
        a[1] -= aux1 * ei / ei1;
        a[2] -= aux2 * ei / ei1;
    } else {
        a[0] -= aux0;
        a[1] -= aux1;
        a[2] -= aux2;
    }
    aux0 = aux1 = aux2 = 0.0;
    // exterior massive particles
    for (int q = 1; q < nm; q++) {
        const int j = S.mlist[q];
        if (massive ? q <= p : j <= i) continue;
        const double t0 = S.mx[q][0] - xi[0], t1 = S.mx[q][1] - xi[1], t2 = S.mx[q][2] - xi[2];
        const double cc = cube_plus(t0, t1, t2, eps3);
        aux0 += G * S.mm[q] * t0 / cc;
        aux1 += G * S.mm[q] * t1 / cc;
        aux2 += G * S.mm[q] * t2 / cc;
    }
    a[0] += aux0;
    a[1] += aux1;
    a[2] += aux2;
    aux0 = aux1 = aux2 = 0.0;
    // pairs straddling i: interior j (ordinal from 0), exterior k; the sum
    // depends only on where i sits among the massive particles
    if (massive) {
        aux0 = S.strad_m[p][0];
        aux1 = S.strad_m[p][1];
        aux2 = S.strad_m[p][2];
    } else {
        aux0 = S.strad_l[ginside][0];
        aux1 = S.strad_l[ginside][1];
        aux2 = S.strad_l[ginside][2];
    }
    a[0] -= aux0 / ei1;
    a[1] -= aux1 / ei1;
    a[2] -= aux2 / ei1;
    jv[3 * i] += a[0] * dt;
    jv[3 * i + 1] += a[1] * dt;
    jv[3 * i + 2] += a[2] * dt;
    if (jv_sync) {
        // the kicked velocity taken back half a step for a snapshot
        // (src/integrator_whfast.c:346-351)
        const double hdt = -0.5 * dt;
        jv_sync[3 * i] = jv[3 * i] + a[0] * hdt;
        jv_sync[3 * i + 1] = jv[3 * i + 1] + a[1] * hdt;
        jv_sync[3 * i + 2] = jv[3 * i + 2] + a[2] * hdt;
    }
}

// The same with the positions read from global memory.
__device__ void accel_kick(const WfShared &S, int i, bool massive, int p, const double *x,
                           const double *jx, const double *m, double G, double eps3,
                           double *jv, double dt, double *jv_sync)
{
    int g = 0;  // massive particles inside i
    if (!massive)
        while (g < S.nm && S.mlist[g] < i) g++;
    accel_kick_v(S, i, massive, p, &x[3 * i], x, &jx[3 * i], g, m[0], G, eps3, jv, dt, jv_sync);
}

__device__ void load_massive_positions(WfShared &S, const double *x)
{
    for (int q = threadIdx.x; q < S.nm; q += WF_BLOCK) {
        const int j = S.mlist[q];
        S.mx[q][0] = x[3 * j];
        S.mx[q][1] = x[3 * j + 1];
        S.mx[q][2] = x[3 * j + 2];
    }
    __syncthreads();
}

// Straddle sums in the reference's order (interior ordinal ascending, then
// exterior ordinal ascending): for a massless particle with g massive
// particles inside it, pairs (q < g <= r); for massive particle p, pairs
// (q < p < r) (src/integrator_whfast.c:930-949, 1099-1119, 1222-1250).
__device__ void straddle_sums(WfShared &S, double G, double eps3)
{
    const int nm = S.nm;
    if (nm <= WF_PAIR_NM) {
        // every pair term once, in parallel; each sum below then adds its
        // terms in the same order as the loop further down
        const int np = nm * (nm - 1) / 2;
        for (int k = threadIdx.x; k < np; k += WF_BLOCK) {
            int q = 0, rem = k;
            while (rem >= nm - 1 - q) {
                rem -= nm - 1 - q;
                q++;
            }
            const int r = q + 1 + rem;
            const double t0 = S.mx[r][0] - S.mx[q][0], t1 = S.mx[r][1] - S.mx[q][1],
                         t2 = S.mx[r][2] - S.mx[q][2];
            const double cc = cube_plus(t0, t1, t2, eps3);
            S.pterm[k][0] = G * S.mm[q] * S.mm[r] * t0 / cc;
            S.pterm[k][1] = G * S.mm[q] * S.mm[r] * t1 / cc;
            S.pterm[k][2] = G * S.mm[q] * S.mm[r] * t2 / cc;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < 2 * nm; t += WF_BLOCK) {
            const bool massive = t > nm;
            const int g = massive ? t - nm : t;
            double a0 = 0.0, a1 = 0.0, a2 = 0.0;
            for (int q = 0; q < g; q++) {
                const int base = q * (2 * nm - q - 1) / 2 - q - 1;  // + r: pair (q, r)
                for (int r = massive ? g + 1 : g; r < nm; r++) {
                    a0 += S.pterm[base + r][0];
                    a1 += S.pterm[base + r][1];
                    a2 += S.pterm[base + r][2];
                }
            }
            double *dst = massive ? S.strad_m[g] : S.strad_l[g];
            dst[0] = a0;
            dst[1] = a1;
            dst[2] = a2;
        }
        __syncthreads();
        return;
    }
    // t = 0..nm: gap g = t; t = nm+1..2nm-1: massive ordinal p = t - nm
    for (int t = threadIdx.x; t < 2 * nm; t += WF_BLOCK) {
        const bool massive = t > nm;
        const int g = massive ? t - nm : t;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int q = 0; q < g; q++)
            for (int r = q + 1; r < nm; r++) {
                if (massive ? r <= g : r < g) continue;
                const double t0 = S.mx[r][0] - S.mx[q][0], t1 = S.mx[r][1] - S.mx[q][1],
                             t2 = S.mx[r][2] - S.mx[q][2];
                const double cc = cube_plus(t0, t1, t2, eps3);
                a0 += G * S.mm[q] * S.mm[r] * t0 / cc;
                a1 += G * S.mm[q] * S.mm[r] * t1 / cc;
                a2 += G * S.mm[q] * S.mm[r] * t2 / cc;
            }
        double *dst = massive ? S.strad_m[g] : S.strad_l[g];
        dst[0] = a0;
        dst[1] = a1;
        dst[2] = a2;
    }
    __syncthreads();
}

// jv_sync != NULL: also the half-step-back velocities of a snapshot.
__device__ void accel_kick_all(WfShared &S, int n, const double *x, const double *jx,
                               const double *m, double G, double eps3, double *jv, double dt,
                               int method, double *jv_sync = nullptr)
{
    load_massive_positions(S, x);
    straddle_sums(S, G, eps3);
    // massive particles (ordinal >= 1), one per thread
    for (int p = 1 + static_cast<int>(threadIdx.x); p < S.nm; p += WF_BLOCK)
        accel_kick(S, S.mlist[p], true, p, x, jx, m, G, eps3, jv, dt, jv_sync);
    if (method == 2) {
        for (int i = 1 + threadIdx.x; i < n; i += WF_BLOCK)
            if (m[i] == 0.0) accel_kick(S, i, false, -1, x, jx, m, G, eps3, jv, dt, jv_sync);
    }
    __syncthreads();
}

// Positions from the quotients fused with the kick: the massive particles'
// positions go to LDS, every massless particle forms its position in
// registers; global positions are written only when `write_x` (they are
// an output, not an input, of the next step).  Requires no component marked
// for the serial recurrence.
__device__ void kick_from_quotients(WfShared &S, const WfScratch &X, int n, double *x,
                                    const double *jx, double G, double eps3, double *jv, double dt,
                                    int method, bool write_x, double *jv_sync)
{
    const int nm = S.nm < WF_MAX_MASSIVE ? S.nm : WF_MAX_MASSIVE;
    for (int q = threadIdx.x; q < nm; q += WF_BLOCK) {
        const int j = S.mlist[q];
        for (int d = 0; d < 3; d++) {
            const double v = q == 0 ? X.j.q0[d] : X.j.mjq[q][d] + X.j.mq[q][d];
            S.mx[q][d] = v;
            if (write_x) x[3 * j + d] = v;
        }
    }
    __syncthreads();
    straddle_sums(S, G, eps3);
    const double m0 = S.mm[0];
    for (int p = 1 + static_cast<int>(threadIdx.x); p < nm; p += WF_BLOCK) {
        const int i = S.mlist[p];
        accel_kick_v(S, i, true, p, S.mx[p], S.mx[0], &jx[3 * i], 0, m0, G, eps3, jv, dt,
                     jv_sync);
    }
    if (method == 2) {
        MassiveCursor mc;
        mc.init(S, nm);
        for (int i = 1 + threadIdx.x; i < n; i += WF_BLOCK) {
            const int g = mc.at(S, nm, i);
            if (S.mlist[g] == i) continue;
            const double jxi[3] = {jx[3 * i], jx[3 * i + 1], jx[3 * i + 2]};
            double xi[3];
            for (int d = 0; d < 3; d++) xi[d] = jxi[d] + run_quotient(S, X, nm, n, g, i, d);
            if (write_x) {
                x[3 * i] = xi[0];
                x[3 * i + 1] = xi[1];
                x[3 * i + 2] = xi[2];
            }
            accel_kick_v(S, i, false, -1, xi, S.mx[0], jxi, g + 1, m0, G, eps3, jv, dt,
                         jv_sync);
        }
    }
    __syncthreads();
}

// Remove flagged particles preserving order (src/system.c:444-528 and the
// Jacobi / eta arrays in src/integrator_whfast.c:624-670).
__device__ int compact(WfShared &S, int n, const int *flag, double *jx, double *jv, double *m,
                       int *ids, double *jx2, double *jv2, double *m2, int *ids2)
{
    int o = 0;
    for (int base = 0; base < n; base += WF_BLOCK) {
        const int i = base + threadIdx.x;
        const int keep = (i < n && !flag[i]) ? 1 : 0;
        int tot;
        const int pos = block_scan_excl(keep, &tot, S.scan_buf);
        if (keep) {
            const int d = o + pos;
            jx2[3 * d] = jx[3 * i];
            jx2[3 * d + 1] = jx[3 * i + 1];
            jx2[3 * d + 2] = jx[3 * i + 2];
            jv2[3 * d] = jv[3 * i];
            jv2[3 * d + 1] = jv[3 * i + 1];
            jv2[3 * d + 2] = jv[3 * i + 2];
            m2[d] = m[i];
            ids2[d] = ids[i];
        }
        o += tot;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < o; i += WF_BLOCK) {
        jx[3 * i] = jx2[3 * i];
        jx[3 * i + 1] = jx2[3 * i + 1];
        jx[3 * i + 2] = jx2[3 * i + 2];
        jv[3 * i] = jv2[3 * i];
        jv[3 * i + 1] = jv2[3 * i + 1];
        jv[3 * i + 2] = jv2[3 * i + 2];
        m[i] = m2[i];
        ids[i] = ids2[i];
    }
    __syncthreads();
    return o;
}

// Initialisation (src/integrator_whfast.c:241-273): sort by Cartesian
// distance, eta, Cartesian -> Jacobi, accelerations, half kick.
__global__ __launch_bounds__(WF_BLOCK) void k_whfast_init(WfArgs A)
{
    __shared__ WfShared S;
    const WfSys sy = A.sys[blockIdx.x];
    const long long o = sy.off;
    int n = A.count[blockIdx.x];
    double *x = A.x + 3 * o, *v = A.v + 3 * o, *jx = A.jx + 3 * o, *jv = A.jv + 3 * o;
    double *m = A.m + o, *m2 = A.m2 + o;
    int *ids = A.ids + o, *ids2 = A.ids2 + o;
    sort_by_distance(S, n, x, v, m, ids, A.jx2 + 3 * o, A.jv2 + 3 * o, m2, ids2);
    massive_list(S, n, m, A.method);
    fill_eta(S, n);
    if (threadIdx.x < 3) c2j_component(n, threadIdx.x, x, jx, m, S);
    else if (threadIdx.x >= 64 && threadIdx.x < 67) c2j_component(n, threadIdx.x - 64, v, jv, m, S);
    __syncthreads();
    const double eps3 = A.eps * A.eps * A.eps;
    accel_kick_all(S, n, x, jx, m, A.G, eps3, jv, 0.5 * A.dt, A.method);
}

// `nsteps` steps of the main loop (src/integrator_whfast.c:290-378).
__global__ __launch_bounds__(WF_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_whfast_steps(WfArgs A)
{
    __shared__ WfShared S;
    __shared__ WfScratch X;
    int *flags = X.flags;
    const WfSys sy = A.sys[blockIdx.x];
    const long long o = sy.off;
    int n = A.count[blockIdx.x];
    double *x = A.x + 3 * o;
    // the state (jx, jv, m, ids) alternates between the two buffer sets: the
    // drift reads each particle from its old place in one set and writes it
    // to its sorted place in the other, so the sort's permutation costs no
    // pass of its own; the other set is scratch for the rest of the step
    double *jx = A.jx + 3 * o, *jv = A.jv + 3 * o, *jx2 = A.jx2 + 3 * o, *jv2 = A.jv2 + 3 * o;
    double *m = A.m + o, *m2 = A.m2 + o;
    int *ids = A.ids + o, *ids2 = A.ids2 + o;
    const double eps3 = A.eps * A.eps * A.eps;
    for (long long s = A.step0; s < A.step0 + A.nsteps; s++) {
        const double dt = (s == A.total_steps - 1) ? A.dt_last : A.dt;
        const bool last = s == A.step0 + A.nsteps - 1;
        sort_keys(S, n, jx, ids);
        // while no massive particle changes place, the massive list, masses
        // and cumulative masses stand
        int moved = s == A.step0 ? 1 : 0;
        const int nmc = S.nm < WF_MAX_MASSIVE ? S.nm : WF_MAX_MASSIVE;
        if (!moved)
            for (int q = threadIdx.x; q < nmc; q += WF_BLOCK) {
                const int j = S.mlist[q];
                if (S.idx[j] != j) moved = 1;
            }
        if (__syncthreads_or(moved)) massive_list(S, n, m, A.method, S.idx);
        fill_eta(S, n);
        if (threadIdx.x == 0) {
            S.any_remove = 0;
            S.bad = 0;
        }
        for (int i = threadIdx.x; i < n; i += WF_BLOCK) flags[i] = 0;
        __syncthreads();
        const double gm0 = A.G * m[S.idx[0]];
        int bad = 0;
        // drift (sorted place i >= 1), reading the old place S.idx[i] of jx,
        // jv, m, ids and writing place i of the other set
        for (int i = threadIdx.x; i < n; i += WF_BLOCK) {
            const int from = S.idx[i];
            double q[3] = {jx[3 * from], jx[3 * from + 1], jx[3 * from + 2]};
            double v[3] = {jv[3 * from], jv[3 * from + 1], jv[3 * from + 2]};
            const double mi = m[from];
            const int id = ids[from];
            if (i >= 1) {
                const double gm = gm0 * S.key[i] / S.key[i - 1];
                const int f = kepler_drift(q, v, gm, dt);
                if (f & 2) atomicOr(&A.status[blockIdx.x], 1);
                if ((f & 1) && A.remove_invalid) {
                    flags[i] = 1;
                    S.any_remove = 1;
                }
                for (int d = 0; d < 3; d++)
                    if (!isfinite(q[d])) bad |= 1 << d;
            }
            for (int d = 0; d < 3; d++) {
                jx2[3 * i + d] = q[d];
                jv2[3 * i + d] = v[d];
            }
            m2[i] = mi;
            ids2[i] = id;
        }
        if (bad) atomicOr(&S.bad, bad);
        __syncthreads();
        {
            double *t = jx; jx = jx2; jx2 = t;
            t = jv; jv = jv2; jv2 = t;
            t = m; m = m2; m2 = t;
            int *u = ids; ids = ids2; ids2 = u;
        }
        // (S.bad is reused for the velocities below)
        const int pos_bad = S.bad;
        if (S.any_remove) {
            n = compact(S, n, flags, jx, jv, m, ids, jx2, jv2, m2, ids2);
            massive_list(S, n, m, A.method);
            fill_eta(S, n);
        }
        __syncthreads();  // the removal flags share their LDS with the quotients
        // Cartesian velocities only matter for the state returned to the
        // caller: rebuild them (from the drifted, not yet kicked, Jacobi
        // velocities, as the reference's last jacobi_to_cartesian does) on
        // the last step of each launch
        // A launch that ends on a snapshot step instead returns the
        // velocities the snapshot holds: the kicked Jacobi velocities taken
        // back half a step (in jv2, free until the next sort), converted
        // after the kick.  Either way A.v is what the reference leaves in
        // system->v at that point.
        const bool snap = last && A.snapshot_last;
        double *jv_sync = snap ? jv2 : nullptr;
        if (last && !snap) j2c_block(n, jv, A.v + 3 * o, m, S, X);
        // (pos_bad covers every drifted particle: a superset of the runs,
        // and compaction only removes particles)
        j2c_quotients(n, jx, S, X, pos_bad);
        if (X.j.fallback[0] | X.j.fallback[1] | X.j.fallback[2]) {
            j2c_fill(n, jx, x, m, S, X);
            accel_kick_all(S, n, x, jx, m, A.G, eps3, jv, dt, A.method, jv_sync);
        } else {
            kick_from_quotients(S, X, n, x, jx, A.G, eps3, jv, dt, A.method, last, jv_sync);
        }
        if (snap) {
            // particle 0 gets no acceleration: jv + 0 * (-dt / 2) == jv
            if (threadIdx.x < 3) jv2[threadIdx.x] = jv[threadIdx.x];
            __syncthreads();
            j2c_block(n, jv2, A.v + 3 * o, m, S, X);
        }
    }
    // an odd number of steps leaves the state in the second set: back to the first
    if (jx != A.jx + 3 * o) {
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += WF_BLOCK) {
            for (int d = 0; d < 3; d++) {
                jx2[3 * i + d] = jx[3 * i + d];
                jv2[3 * i + d] = jv[3 * i + d];
            }
            m2[i] = m[i];
            ids2[i] = ids[i];
        }
    }
    if (threadIdx.x == 0) A.count[blockIdx.x] = n;
}

// ---------------------------------------------------------------- host
ErrorStatus whfast_ensemble(int num_systems, const int64 *offsets, int32 *ids, double *x, double *v,
                            double *m, double G, int method, double eps, double dt, double tf,
                            bool remove_invalid, int *out_counts, int64 *steps_done,
                            const bool *is_exit, SnapshotWriter *snaps,
                            const SimulationStatus *status0, const OutputParam *op)
{
    if (snaps && snaps->on() && num_systems != 1)
        return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR, "snapshots are written for one system only");
    if (num_systems <= 0 || !offsets || !ids || !x || !v || !m || !out_counts)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument or no systems");
    if (method != ACCELERATION_METHOD_PAIRWISE && method != ACCELERATION_METHOD_MASSLESS)
        return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR,
                                "Invalid acceleration method for WHFast integrator. Only pairwise "
                                "and massless methods are supported.");
    if (dt <= 0.0) return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "dt must be positive. Got: %g", dt);
    std::vector<WfSys> sys(num_systems);
    for (int s = 0; s < num_systems; s++) {
        const long long n = offsets[s + 1] - offsets[s];
        if (n <= 0 || n > WF_MAX_N)
            return WRAP_RAISE_ERROR_FMT(GRAV_NOT_IMPLEMENTED_ERROR,
                                        "System %d has %lld particles; the GPU WHFast supports "
                                        "1..%d per system",
                                        s, n, WF_MAX_N);
        int nm = 0;
        for (long long i = offsets[s]; i < offsets[s + 1]; i++)
            nm += (method == ACCELERATION_METHOD_PAIRWISE || m[i] != 0.0) ? 1 : 0;
        if (nm > WF_MAX_MASSIVE)
            return WRAP_RAISE_ERROR_FMT(GRAV_NOT_IMPLEMENTED_ERROR,
                                        "System %d has %d massive particles; the GPU WHFast "
                                        "supports up to %d",
                                        s, nm, WF_MAX_MASSIVE);
        sys[s].off = offsets[s];
        sys[s].cap = static_cast<int>(n);
    }
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    const long long P = offsets[num_systems] - offsets[0];
    const long long base = offsets[0];
    for (auto &s : sys) s.off -= base;
    // workgroup b integrates system order[b]: largest systems first, so the
    // uneven systems pack the CUs with a short tail (per-workgroup counts and
    // status follow the same order and are mapped back below)
    std::vector<int> order(num_systems);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return sys[a].cap > sys[b].cap; });
    {
        std::vector<WfSys> by(num_systems);
        for (int b = 0; b < num_systems; b++) by[b] = sys[order[b]];
        sys.swap(by);
    }
    // one allocation: 7 x 3P doubles (x, v, jx, jv, jx2, jv2) + 2P doubles (m, m2)
    // + 2P ints (ids, ids2) + per-system arrays
    const size_t b3 = sizeof(double) * 3 * P;
    // (each of the 13 sub-buffers below is rounded up to 256 bytes)
    const size_t bytes = 6 * b3 + 2 * sizeof(double) * P + 2 * sizeof(int) * P +
                         sizeof(WfSys) * num_systems + 2 * sizeof(int) * num_systems + 13 * 256;
    char *buf = nullptr;
    GS_HIP(hipMalloc(&buf, bytes));
    WfArgs A;
    char *p = buf;
    auto take = [&](size_t b) {
        char *r = p;
        p += (b + 255) & ~static_cast<size_t>(255);
        return r;
    };
    // p never passes buf + bytes: 13 takes of their size rounded up to 256
    A.x = reinterpret_cast<double *>(take(b3));
    A.v = reinterpret_cast<double *>(take(b3));
    A.jx = reinterpret_cast<double *>(take(b3));
    A.jv = reinterpret_cast<double *>(take(b3));
    A.jx2 = reinterpret_cast<double *>(take(b3));
    A.jv2 = reinterpret_cast<double *>(take(b3));
    A.m = reinterpret_cast<double *>(take(sizeof(double) * P));
    A.m2 = reinterpret_cast<double *>(take(sizeof(double) * P));
    A.ids = reinterpret_cast<int *>(take(sizeof(int) * P));
    A.ids2 = reinterpret_cast<int *>(take(sizeof(int) * P));
    WfSys *dsys = reinterpret_cast<WfSys *>(take(sizeof(WfSys) * num_systems));
    A.count = reinterpret_cast<int *>(take(sizeof(int) * num_systems));
    A.status = reinterpret_cast<int *>(take(sizeof(int) * num_systems));
    A.sys = dsys;
    std::vector<int> cnt(num_systems);
    for (int s = 0; s < num_systems; s++) cnt[s] = sys[s].cap;
    ErrorStatus st = make_success_error_status();
    hipError_t e = hipMemcpyAsync(A.x, x + 3 * base, b3, hipMemcpyHostToDevice, c->stream);
    if (!e) e = hipMemcpyAsync(A.v, v + 3 * base, b3, hipMemcpyHostToDevice, c->stream);
    if (!e) e = hipMemcpyAsync(A.m, m + base, sizeof(double) * P, hipMemcpyHostToDevice, c->stream);
    if (!e) e = hipMemcpyAsync(A.ids, ids + base, sizeof(int) * P, hipMemcpyHostToDevice, c->stream);
    if (!e) e = hipMemcpyAsync(dsys, sys.data(), sizeof(WfSys) * num_systems, hipMemcpyHostToDevice, c->stream);
    if (!e) e = hipMemcpyAsync(A.count, cnt.data(), sizeof(int) * num_systems, hipMemcpyHostToDevice, c->stream);
    if (!e) e = hipMemsetAsync(A.status, 0, sizeof(int) * num_systems, c->stream);
    if (!e) e = hipMemsetAsync(A.jx, 0, b3, c->stream);
    A.G = G;
    A.eps = eps;
    A.dt = dt;
    A.method = method;
    A.remove_invalid = remove_invalid ? 1 : 0;
    // dt clamp of the last step (src/integrator_whfast.c:292-296)
    const long long total = static_cast<long long>(ceil(tf / dt));
    A.total_steps = total;
    A.dt_last = dt;
    if (total > 0) {
        const double t_last = static_cast<double>(total - 1) * dt;
        if (t_last + dt > tf) A.dt_last = tf - t_last;
    }
    A.step0 = 0;
    A.nsteps = 0;
    A.snapshot_last = 0;
    if (!e) {
        const int rec = timing_begin(c, "whfast_init");
        hipLaunchKernelGGL(k_whfast_init, dim3(num_systems), dim3(WF_BLOCK), 0, c->stream, A);
        timing_end(c, rec);
        e = hipGetLastError();
    }
    const bool out = snaps && snaps->on();
    // initial snapshot of the distance-sorted state (src/integrator_whfast.c:240-262)
    if (!e && out && op->output_initial)
        st = snaps->write_device_particles(A.x, A.v, A.m, A.ids, A.count, status0->t,
                                           status0->dt);
    const long long chunk = 256;
    long long done = 0;
    while (!e && st.return_code == GRAV_SUCCESS && done < total) {
        A.step0 = done;
        A.nsteps = std::min(chunk, total - done);
        A.snapshot_last = 0;
        // a launch ends at the first step whose time t = steps dt reaches the
        // next output time (src/integrator_whfast.c:343-366)
        double t_end = 0.0, dt_end = dt;
        if (out)
            for (long long j = 1; j <= A.nsteps; j++) {
                const long long k = done + j;
                const double dtk = k == total ? A.dt_last : dt;
                if (snaps->due(static_cast<double>(k) * dtk)) {
                    A.nsteps = j;
                    A.snapshot_last = 1;
                    t_end = static_cast<double>(k) * dtk;
                    dt_end = dtk;
                    break;
                }
            }
        const int rec = timing_begin(c, "whfast_steps");
        hipLaunchKernelGGL(k_whfast_steps, dim3(num_systems), dim3(WF_BLOCK), 0, c->stream, A);
        timing_end(c, rec);
        e = hipGetLastError();
        done += A.nsteps;
        if (!e && A.snapshot_last)
            st = snaps->write_device_particles(A.x, A.v, A.m, A.ids, A.count, t_end, dt_end);
        if (!e) e = hipStreamSynchronize(c->stream);
        if (is_exit && *is_exit) break;
    }
    if (st.return_code != GRAV_SUCCESS) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(buf);
        return st;
    }
    std::vector<int> stat(num_systems);
    if (!e) e = hipMemcpyAsync(cnt.data(), A.count, sizeof(int) * num_systems, hipMemcpyDeviceToHost, c->stream);
    if (!e) e = hipMemcpyAsync(stat.data(), A.status, sizeof(int) * num_systems, hipMemcpyDeviceToHost, c->stream);
    if (!e) e = hipMemcpyAsync(x + 3 * base, A.x, b3, hipMemcpyDeviceToHost, c->stream);
    if (!e) e = hipMemcpyAsync(v + 3 * base, A.v, b3, hipMemcpyDeviceToHost, c->stream);
    if (!e) e = hipMemcpyAsync(m + base, A.m, sizeof(double) * P, hipMemcpyDeviceToHost, c->stream);
    if (!e) e = hipMemcpyAsync(ids + base, A.ids, sizeof(int) * P, hipMemcpyDeviceToHost, c->stream);
    if (!e) e = hipStreamSynchronize(c->stream);
    (void)hipFree(buf);
    if (e) return hip_error(e, "GPU WHFast", __FILE__, __LINE__, __func__);
    for (int b = 0; b < num_systems; b++) out_counts[order[b]] = cnt[b];
    if (steps_done) *steps_done = done;
    return st;
}

ErrorStatus whfast_gpu(System *system, IntegratorParam *ip, AccelerationParam *ap,
                       OutputParam *op, SimulationStatus *status, Settings *settings, double tf)
{
    const int64 off[2] = {0, system->num_particles};
    int count = 0;
    int64 steps = 0;
    Context *c;
    GS_TRY(context(&c));
    SnapshotWriter snaps;
    GS_TRY(snaps.begin(c, op, system));
    GS_TRY(whfast_ensemble(1, off, system->particle_ids, system->x, system->v, system->m,
                           system->G, ap->method, ap->softening_length, ip->dt, tf,
                           ip->whfast_remove_invalid_particles, &count, &steps,
                           settings->is_exit_ptr, &snaps, status, op));
    GS_TRY(snaps.finish());
    system->num_particles = count;
    status->num_steps = steps;
    const long long total = static_cast<long long>(ceil(tf / ip->dt));
    double dt = ip->dt;
    if (total > 0 && steps == total) {
        const double t_last = static_cast<double>(total - 1) * dt;
        if (t_last + dt > tf) dt = tf - t_last;
    }
    status->dt = dt;
    status->t = static_cast<double>(steps) * dt;
    return make_success_error_status();
}

}  // namespace gs

extern "C" ErrorStatus gs_whfast_ensemble(int num_systems, const int64 *offsets, int32 *ids,
                                          double *x, double *v, double *m, double G, int method,
                                          double eps, double dt, double tf, bool remove_invalid,
                                          int *out_counts)
{
    return gs::whfast_ensemble(num_systems, offsets, ids, x, v, m, G, method, eps, dt, tf,
                               remove_invalid, out_counts, nullptr, nullptr, nullptr, nullptr,
                               nullptr);
}
