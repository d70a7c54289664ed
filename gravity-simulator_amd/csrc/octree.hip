// octree.hip -- Barnes-Hut linear octree on gfx950.
//
// Pipeline (src/linear_octree.c:113-962, src/acceleration_barnes_hut.c:78-248):
//   k_bbox_*       min/max reduction of x (FP64, order independent -> exact)
//   k_morton       63-bit keys with the reference's FP64 quantisation
//                  ((x-c)/w + 0.5, *2^21, truncate, mask) and magic-number
//                  bit spread; contraction off -> bit-exact keys
//   radix sort     stable LSD, 8-bit digits, 8 passes (histogram, scan,
//                  wave-ranked scatter): the same permutation as the
//                  reference's stable 9-bit LSD sort
//   build          leaf capacity <= 16: every level at once from the shared
//                  octal digits of neighbouring keys (count per depth and
//                  tile, one scan, emission in level order); larger leaves:
//                  level-synchronous (children of one depth found by binary
//                  search on the sorted keys, counted, scanned, emitted)
//   multipoles     bottom-up per level; leaves sum their particles in sorted
//                  order, internal nodes sum their children (FP64)
//   k_walk         one wavefront walks for 64 Morton-consecutive targets with
//                  a shared explicit stack of (node, lane mask): each lane
//                  applies the reference opening test to its own particle;
//                  lanes that accept a node add its monopole, the others
//                  descend.  Geometry and sums in FP64, 1/r^3 through FP32
//                  rsqrt.
#include <cmath>
#include <cstring>

#include "gs_internal.h"
#include "octree.h"
#include "scan.h"

#pragma clang fp contract(off)

namespace gs {

constexpr int MAXL = MORTON_MAX_LEVEL;

// ------------------------------------------------------------------ bbox
constexpr int BB_BLOCK = 256;

__global__ __launch_bounds__(BB_BLOCK) void k_bbox_partial(const double *__restrict__ x, int n,
                                                           double *__restrict__ part)
{
    __shared__ double red[6][BB_BLOCK];
    double lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
        lo[d] = x[d];
        hi[d] = x[d];
    }
    for (int i = blockIdx.x * BB_BLOCK + threadIdx.x; i < n; i += gridDim.x * BB_BLOCK)
        for (int d = 0; d < 3; d++) {
            lo[d] = fmin(lo[d], x[3 * i + d]);
            hi[d] = fmax(hi[d], x[3 * i + d]);
        }
    for (int d = 0; d < 3; d++) {
        red[d][threadIdx.x] = lo[d];
        red[3 + d][threadIdx.x] = hi[d];
    }
    __syncthreads();
    for (int s = BB_BLOCK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int d = 0; d < 3; d++) {
                red[d][threadIdx.x] = fmin(red[d][threadIdx.x], red[d][threadIdx.x + s]);
                red[3 + d][threadIdx.x] = fmax(red[3 + d][threadIdx.x], red[3 + d][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// out[0..2] = centre, out[3] = width (src/linear_octree.c:138-145); one
// workgroup of BB_BLOCK threads folds the partial boxes
__global__ __launch_bounds__(BB_BLOCK) void k_bbox_final(const double *__restrict__ part,
                                                         int nparts, double *__restrict__ out)
{
    __shared__ double red[6][BB_BLOCK];
    double lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
        lo[d] = part[d];
        hi[d] = part[3 + d];
    }
    for (int p = threadIdx.x; p < nparts; p += BB_BLOCK)
        for (int d = 0; d < 3; d++) {
            lo[d] = fmin(lo[d], part[6 * p + d]);
            hi[d] = fmax(hi[d], part[6 * p + 3 + d]);
        }
    for (int d = 0; d < 3; d++) {
        red[d][threadIdx.x] = lo[d];
        red[3 + d][threadIdx.x] = hi[d];
    }
    __syncthreads();
    for (int s = BB_BLOCK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int d = 0; d < 3; d++) {
                red[d][threadIdx.x] = fmin(red[d][threadIdx.x], red[d][threadIdx.x + s]);
                red[3 + d][threadIdx.x] = fmax(red[3 + d][threadIdx.x], red[3 + d][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int d = 0; d < 3; d++) {
            lo[d] = red[d][0];
            hi[d] = red[3 + d][0];
        }
        for (int d = 0; d < 3; d++) out[d] = (hi[d] + lo[d]) / 2.0;
        out[3] = fmax(fmax(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
    }
}

// ------------------------------------------------------------------ Morton
__device__ __forceinline__ long long spread21(long long v)
{
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffLL;
    v = (v | v << 16) & 0x1f0000ff0000ffLL;
    v = (v | v << 8) & 0x100f00f00f00f00fLL;
    v = (v | v << 4) & 0x10c30c30c30c30c3LL;
    v = (v | v << 2) & 0x1249249249249249LL;
    return v;
}

__device__ __forceinline__ long long quantise(double xv, double c, double w)
{
    const double u = (xv - c) / w + 0.5;
    const double s = u * 2097152.0;
    // C's double -> int64 conversion truncates toward zero; out-of-range and
    // NaN inputs (zero-width box) map to 0 here instead of being undefined
    if (!(s > -9.2e18 && s < 9.2e18)) return 0;
    return static_cast<long long>(s);
}

// With packed != NULL it also writes (x, y, z, m) per particle as one
// 32-byte record: the gather into sorted order then reads one aligned record
// per particle instead of a 24-byte and an 8-byte piece of two arrays.
__global__ void k_morton(const double *__restrict__ x, int n, const double *__restrict__ box,
                         long long *__restrict__ keys, int *__restrict__ idx,
                         const double *__restrict__ m, double4 *__restrict__ packed)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double w = box[3];
    const double x0 = x[3 * i + 0], x1 = x[3 * i + 1], x2 = x[3 * i + 2];
    const long long nx = spread21(quantise(x0, box[0], w));
    const long long ny = spread21(quantise(x1, box[1], w));
    const long long nz = spread21(quantise(x2, box[2], w));
    keys[i] = nx | (ny << 1) | (nz << 2);
    idx[i] = i;
    if (packed) packed[i] = make_double4(x0, x1, x2, m[i]);
}

// ------------------------------------------------------------------ radix sort
constexpr int RS_BLOCK = 256;
constexpr int RS_WAVES = RS_BLOCK / 64;
constexpr int RS_ROUNDS = 8;                          // keys per lane
constexpr int RS_TILE = RS_BLOCK * RS_ROUNDS;         // 2048 keys per tile
constexpr int RS_BITS = 8;
constexpr int RS_BINS = 1 << RS_BITS;
constexpr int RS_PASSES = 8;                          // 64 bits (bit 63 is always 0)

__global__ __launch_bounds__(RS_BLOCK) void k_rs_hist(const long long *__restrict__ keys, int n,
                                                      int shift, unsigned *__restrict__ hist,
                                                      int ntiles)
{
    __shared__ unsigned h[RS_BINS];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * RS_TILE;
    for (int k = 0; k < RS_ROUNDS; k++) {
        const int q = base + k * RS_BLOCK + threadIdx.x;
        if (q < n) atomicAdd(&h[(keys[q] >> shift) & (RS_BINS - 1)], 1u);
    }
    __syncthreads();
    // digit-major layout: all tiles of digit 0, then digit 1, ...
    hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of each digit's row of tile counts (digit-major layout: a
// digit's counts for every tile are contiguous), one workgroup per digit;
// dtot[d] := the digit's total, whose exclusive scan over the digits the
// scatter adds.  A thread takes RS_ROW consecutive tiles per chunk of
// RS_BLOCK * RS_ROW, so one chunk covers up to 2^22 keys.
constexpr int RS_ROW = 8;
__global__ __launch_bounds__(RS_BLOCK) void k_rs_rowscan(unsigned *__restrict__ hist, int ntiles,
                                                         unsigned *__restrict__ dtot)
{
    __shared__ unsigned wsum[RS_WAVES];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    unsigned *row = hist + static_cast<size_t>(blockIdx.x) * ntiles;
    unsigned carry = 0;
    for (int base = 0; base < ntiles; base += RS_BLOCK * RS_ROW) {
        const int q0 = base + threadIdx.x * RS_ROW;
        unsigned v[RS_ROW];
#pragma unroll
        for (int k = 0; k < RS_ROW; k++) v[k] = q0 + k < ntiles ? row[q0 + k] : 0u;
        unsigned mine = 0;
#pragma unroll
        for (int k = 0; k < RS_ROW; k++) mine += v[k];
        unsigned incl = mine;
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned u = __shfl_up(incl, off, 64);
            if (lane >= off) incl += u;
        }
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        unsigned run = carry, all = 0;
        for (int w = 0; w < RS_WAVES; w++) {
            if (w < wave) run += wsum[w];
            all += wsum[w];
        }
        run += incl - mine;
#pragma unroll
        for (int k = 0; k < RS_ROW; k++) {
            if (q0 + k < ntiles) row[q0 + k] = run;
            run += v[k];
        }
        carry += all;
        __syncthreads();  // wsum is rewritten by the next chunk
    }
    if (threadIdx.x == 0) dtot[blockIdx.x] = carry;
}

// Stable scatter: wave w owns tile positions [w*512, (w+1)*512), visited in
// 8 rounds of 64 consecutive keys; within a round, peers with equal digits
// are found with 8 ballots and ranked by lane order.  The tile is then
// reordered by digit in LDS, so that consecutive threads write consecutive
// addresses of each digit's run instead of scattering single keys.
static_assert(RS_BLOCK == RS_BINS, "one thread per digit for the tile offsets");
__global__ __launch_bounds__(RS_BLOCK) void k_rs_scatter(const long long *__restrict__ kin,
                                                         const int *__restrict__ iin, int n,
                                                         int shift,
                                                         const unsigned *__restrict__ offs,
                                                         const unsigned *__restrict__ dtot,
                                                         int ntiles,
                                                         long long *__restrict__ kout,
                                                         int *__restrict__ iout)
{
    __shared__ unsigned wcount[RS_WAVES][RS_BINS];
    __shared__ unsigned toff[RS_BINS];   // first tile position of each digit
    __shared__ unsigned goff[RS_BINS];   // first output position of each digit
    __shared__ unsigned wsum[RS_WAVES];
    __shared__ unsigned dsum[RS_WAVES];
    __shared__ long long skey[RS_TILE];
    __shared__ int sval[RS_TILE];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    for (int b = threadIdx.x; b < RS_WAVES * RS_BINS; b += RS_BLOCK)
        (&wcount[0][0])[b] = 0;
    __syncthreads();

    volatile unsigned *cnt = wcount[wave];
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
    const int base = blockIdx.x * RS_TILE + wave * (RS_ROUNDS * 64);
    long long key[RS_ROUNDS];
    int val[RS_ROUNDS];
    unsigned rank[RS_ROUNDS];
    int dig[RS_ROUNDS];
    for (int r = 0; r < RS_ROUNDS; r++) {
        const int q = base + r * 64 + lane;
        const bool ok = q < n;
        key[r] = ok ? kin[q] : 0;
        val[r] = ok ? iin[q] : 0;
        const int d = ok ? static_cast<int>((key[r] >> shift) & (RS_BINS - 1)) : -1;
        dig[r] = d;
        unsigned long long peers = __ballot(ok);
        for (int b = 0; b < RS_BITS; b++) {
            const bool bit = ok && ((d >> b) & 1);
            const unsigned long long bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        if (!ok) peers = 0;
        const unsigned before = ok ? cnt[d] : 0u;
        rank[r] = before + static_cast<unsigned>(__popcll(peers & lt_mask));
        __builtin_amdgcn_wave_barrier();
        // the lowest lane of each peer group advances the wave's count
        if (ok && (peers & lt_mask) == 0) cnt[d] = before + static_cast<unsigned>(__popcll(peers));
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // tile offsets of the digits: exclusive scan of the tile histogram; the
    // digits' first output positions: exclusive scan of the digit totals
    // plus this tile's offset within its digit's row
    {
        const int d = threadIdx.x;
        unsigned tot = 0;
        for (int w = 0; w < RS_WAVES; w++) tot += wcount[w][d];
        const unsigned dt = dtot[d];
        unsigned incl = tot, dincl = dt;
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned u = __shfl_up(incl, off, 64);
            const unsigned du = __shfl_up(dincl, off, 64);
            if (lane >= off) {
                incl += u;
                dincl += du;
            }
        }
        if (lane == 63) {
            wsum[wave] = incl;
            dsum[wave] = dincl;
        }
        const unsigned row = offs[d * ntiles + blockIdx.x];
        __syncthreads();
        unsigned add = 0, dadd = 0;
        for (int w = 0; w < wave; w++) {
            add += wsum[w];
            dadd += dsum[w];
        }
        toff[d] = add + incl - tot;
        goff[d] = dadd + dincl - dt + row;
    }
    __syncthreads();
    for (int r = 0; r < RS_ROUNDS; r++) {
        const int d = dig[r];
        if (d < 0) continue;
        unsigned lp = toff[d] + rank[r];
        for (int w = 0; w < wave; w++) lp += wcount[w][d];
        skey[lp] = key[r];
        sval[lp] = val[r];
    }
    __syncthreads();
    const int tile_n = min(RS_TILE, n - blockIdx.x * RS_TILE);
    for (int p = threadIdx.x; p < tile_n; p += RS_BLOCK) {
        const long long k = skey[p];
        const int d = static_cast<int>((k >> shift) & (RS_BINS - 1));
        const unsigned dst = goff[d] + (static_cast<unsigned>(p) - toff[d]);
        if (dst < static_cast<unsigned>(n)) {
            kout[dst] = k;
            iout[dst] = sval[p];
        }
    }
}

// ------------------------------------------------------------------ gather
__global__ void k_gather_particles(const double4 *__restrict__ packed,
                                   const int *__restrict__ sidx, int n,
                                   double4 *__restrict__ P)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    P[s] = packed[sidx[s]];
}

// ------------------------------------------------------------------ build
// Child octant ranges of a node of depth d (children at depth d+1) by binary
// search on its sorted key range (src/linear_octree.c:342-393).  The eight
// searches (first position whose octant digit is > o) are independent and
// branch-free with a common trip count, so their key loads overlap instead
// of forming one chain of dependent loads.
__device__ void octant_bounds(const long long *__restrict__ keys, int first, int count,
                              int child_depth, int bounds[9])
{
    const int shift = 3 * (MAXL - child_depth);
    int base[8];
#pragma unroll
    for (int o = 0; o < 8; o++) base[o] = first;
    bounds[0] = first;
    if (count <= 0) {
#pragma unroll
        for (int o = 0; o < 8; o++) bounds[o + 1] = first;
        return;
    }
    // invariant: the answer for octant o lies in [base[o], base[o] + len]
    int len = count;
    while (len > 1) {
        const int half = len >> 1;
#pragma unroll
        for (int o = 0; o < 8; o++) {
            const int oct = static_cast<int>((keys[base[o] + half - 1] >> shift) & 7);
            base[o] = oct <= o ? base[o] + half : base[o];
        }
        len -= half;
    }
#pragma unroll
    for (int o = 0; o < 8; o++) {
        const int oct = static_cast<int>((keys[base[o]] >> shift) & 7);
        bounds[o + 1] = base[o] + (oct <= o ? 1 : 0);
    }
}

__device__ __forceinline__ bool node_internal(const NodeRec &nd, int leaf_cap)
{
    return nd.depth == 0 || (nd.count > leaf_cap && nd.depth < MAXL);
}

__global__ void k_count_children(NodeRec *__restrict__ nodes, int lb, int le, int leaf_cap,
                                 const long long *__restrict__ keys, unsigned *__restrict__ cnt)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= le - lb) return;
    NodeRec nd = nodes[lb + t];
    unsigned c = 0;
    if (node_internal(nd, leaf_cap)) {
        int b[9];
        octant_bounds(keys, nd.first, nd.count, nd.depth + 1, b);
        for (int o = 0; o < 8; o++) c += b[o + 1] > b[o] ? 1u : 0u;
    }
    cnt[t] = c;
}

__global__ void k_emit_children(NodeRec *__restrict__ nodes, int lb, int le, int leaf_cap, double w2,
                                const long long *__restrict__ keys,
                                const unsigned *__restrict__ offs, int node_cap)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= le - lb) return;
    NodeRec &nd = nodes[lb + t];
    if (!node_internal(nd, leaf_cap)) {
        nd.nchild = 0;
        nd.child0 = -1;
        return;
    }
    int b[9];
    octant_bounds(keys, nd.first, nd.count, nd.depth + 1, b);
    const int c0 = le + static_cast<int>(offs[t]);
    int c = c0;
    for (int o = 0; o < 8; o++) {
        if (b[o + 1] <= b[o]) continue;
        NodeRec ch;
        ch.depth = nd.depth + 1;
        ch.prefix = nd.prefix * 8 + o;
        ch.first = b[o];
        ch.count = b[o + 1] - b[o];
        ch.nchild = 0;
        ch.child0 = -1;
        ch.cx = ch.cy = ch.cz = 0.0;
        // (2w / (2 << d))^2 == w^2 * 4^-d exactly (power-of-two scaling)
        ch.l2 = __builtin_amdgcn_ldexp(w2, -2 * ch.depth);
        ch.fmass = 0.0f;
        if (c < node_cap) nodes[c] = ch;
        c++;
    }
    nd.child0 = c0;
    nd.nchild = c - c0;
}

// Monopole sums of the sorted particles [first, first + count), in order, and
// the centre of mass the walks use (X / M, or the first particle's position
// for a massless node).
__device__ __forceinline__ double4 leaf_monopole(const double4 *__restrict__ P, int first, int count,
                                                 NodeRec &nd)
{
    double M = 0.0, X = 0.0, Y = 0.0, Z = 0.0;
    for (int s = first; s < first + count; s++) {
        const double4 p = P[s];
        M += p.w;
        X += p.w * p.x;
        Y += p.w * p.y;
        Z += p.w * p.z;
    }
    nd.fmass = static_cast<float>(M);
    if (M != 0.0) {
        nd.cx = X / M;
        nd.cy = Y / M;
        nd.cz = Z / M;
    } else {
        const double4 p = P[first];
        nd.cx = p.x;
        nd.cy = p.y;
        nd.cz = p.z;
    }
    return make_double4(M, X, Y, Z);
}

// FP32 second central moments of a leaf's particles about (c0, c1, c2), the
// walks' copy (xx yy zz xy xz yz, 1.5 tr, 0)
__device__ __forceinline__ void leaf_quadf(const double4 *__restrict__ P, int first, int count,
                                           double c0, double c1, double c2, float4 *qf)
{
    float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f, f3 = 0.0f, f4 = 0.0f, f5 = 0.0f;
    for (int s = first; s < first + count; s++) {
        const double4 p = P[s];
        const float d0 = static_cast<float>(p.x - c0), d1 = static_cast<float>(p.y - c1),
                    d2 = static_cast<float>(p.z - c2);
        const float mp = static_cast<float>(p.w);
        f0 = __builtin_fmaf(mp * d0, d0, f0);
        f1 = __builtin_fmaf(mp * d1, d1, f1);
        f2 = __builtin_fmaf(mp * d2, d2, f2);
        f3 = __builtin_fmaf(mp * d0, d1, f3);
        f4 = __builtin_fmaf(mp * d0, d2, f4);
        f5 = __builtin_fmaf(mp * d1, d2, f5);
    }
    qf[0] = make_float4(f0, f1, f2, f3);
    qf[1] = make_float4(f4, f5, 1.5f * (f0 + f1 + f2), 0.0f);
}

// moments: mom[node] = (sum m, sum m x, sum m y, sum m z); with QUAD also
// the second central moments (only the quadrupole, expansion and feature
// walks read them): in FP32 from the children's FP32 moments, or with OCT in
// FP64 (the third moments' shifts use them) and rounded to FP32
// Also marks the group walk's parent nodes (the root and every node of more
// than 64 particles) in gflag, from the record it reads anyway.
template <bool QUAD, bool OCT>
__global__ void k_multipole(NodeRec *__restrict__ nodes, int lb, int le,
                            const double4 *__restrict__ P, double4 *__restrict__ mom,
                            double *__restrict__ quad, float *__restrict__ quadf,
                            double *__restrict__ oct, float *__restrict__ octf,
                            unsigned *__restrict__ gflag)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= le - lb) return;
    const int node = lb + t;
    // the record is read whole and written back whole (neighbouring threads
    // write neighbouring records: full cache lines, no partial-line merges)
    NodeRec nd = nodes[node];
    gflag[node] = (node == 0 || nd.count > 64) ? 1u : 0u;
    double4 mnode;
    if (nd.nchild == 0) {
        mnode = leaf_monopole(P, nd.first, nd.count, nd);
    } else {
        // the children's moments are requested before any is summed (loads
        // of up to 4 at once, clamped to the last child), then added in
        // child order
        double M = 0.0, X = 0.0, Y = 0.0, Z = 0.0;
        const int cl = nd.child0 + nd.nchild - 1;
        for (int c = nd.child0; c <= cl; c += 4) {
            double4 q[4];
#pragma unroll
            for (int k = 0; k < 4; k++) q[k] = mom[c + k < cl ? c + k : cl];
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (c + k <= cl) {
                    M += q[k].x;
                    X += q[k].y;
                    Y += q[k].z;
                    Z += q[k].w;
                }
        }
        mnode = make_double4(M, X, Y, Z);
        nd.fmass = static_cast<float>(M);
        if (M != 0.0) {
            nd.cx = X / M;
            nd.cy = Y / M;
            nd.cz = Z / M;
        } else {
            // massless node: place its (zero) monopole on its first particle
            const double4 p = P[nd.first];
            nd.cx = p.x;
            nd.cy = p.y;
            nd.cz = p.z;
        }
    }
    mom[node] = mnode;
    nodes[node] = nd;
    if (!QUAD) return;
    const double c0 = nd.cx, c1 = nd.cy, c2 = nd.cz;
    float4 *qf = reinterpret_cast<float4 *>(quadf + 8 * static_cast<size_t>(node));
    if (!OCT) {
        // second central moments in FP32 only, the walks' copy (xx yy zz xy
        // xz yz, 1.5 tr): leaves from their particles, internal nodes from
        // their children's FP32 moments shifted to this centre of mass
        // (parallel axes; the offsets formed in FP64)
        if (nd.nchild == 0) {
            leaf_quadf(P, nd.first, nd.count, c0, c1, c2, qf);
            return;
        }
        float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f, f3 = 0.0f, f4 = 0.0f, f5 = 0.0f;
        const float4 *qs = reinterpret_cast<const float4 *>(quadf);
        const int cl = nd.child0 + nd.nchild - 1;
        for (int c = nd.child0; c <= cl; c += 4) {
            double4 mk[4];
            float4 qa[4], qb[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int ck = c + k < cl ? c + k : cl;
                mk[k] = mom[ck];
                qa[k] = qs[2 * static_cast<size_t>(ck)];
                qb[k] = qs[2 * static_cast<size_t>(ck) + 1];
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (c + k > cl) break;
                const double4 mc = mk[k];
                double cc0, cc1, cc2;
                if (mc.x != 0.0) {
                    cc0 = mc.y / mc.x;
                    cc1 = mc.z / mc.x;
                    cc2 = mc.w / mc.x;
                } else {
                    const double4 p = P[nodes[c + k].first];
                    cc0 = p.x;
                    cc1 = p.y;
                    cc2 = p.z;
                }
                const float d0 = static_cast<float>(cc0 - c0), d1 = static_cast<float>(cc1 - c1),
                            d2 = static_cast<float>(cc2 - c2);
                const float Mc = static_cast<float>(mc.x);
                f0 += __builtin_fmaf(Mc * d0, d0, qa[k].x);
                f1 += __builtin_fmaf(Mc * d1, d1, qa[k].y);
                f2 += __builtin_fmaf(Mc * d2, d2, qa[k].z);
                f3 += __builtin_fmaf(Mc * d0, d1, qa[k].w);
                f4 += __builtin_fmaf(Mc * d0, d2, qb[k].x);
                f5 += __builtin_fmaf(Mc * d1, d2, qb[k].y);
            }
        }
        qf[0] = make_float4(f0, f1, f2, f3);
        qf[1] = make_float4(f4, f5, 1.5f * (f0 + f1 + f2), 0.0f);
        return;
    }
    // with third moments: the second central moments Q_ab = sum m (x -
    // com)_a (x - com)_b in FP64 (xx yy zz xy xz yz), which the third
    // moments' shifts need: leaves from their particles, internal nodes from
    // their children shifted to this centre of mass (parallel axes)
    double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0, q4 = 0.0, q5 = 0.0;
    if (nd.nchild == 0) {
        for (int s = nd.first; s < nd.first + nd.count; s++) {
            const double4 p = P[s];
            const double d0 = p.x - c0, d1 = p.y - c1, d2 = p.z - c2;
            q0 += p.w * d0 * d0;
            q1 += p.w * d1 * d1;
            q2 += p.w * d2 * d2;
            q3 += p.w * d0 * d1;
            q4 += p.w * d0 * d2;
            q5 += p.w * d1 * d2;
        }
    } else {
        for (int c = nd.child0; c < nd.child0 + nd.nchild; c++) {
            // the child's centre of mass as its record holds it (X / M, or
            // its first particle when massless), without reading the record
            const double4 mc = mom[c];
            double cc0, cc1, cc2;
            if (mc.x != 0.0) {
                cc0 = mc.y / mc.x;
                cc1 = mc.z / mc.x;
                cc2 = mc.w / mc.x;
            } else {
                const double4 p = P[nodes[c].first];
                cc0 = p.x;
                cc1 = p.y;
                cc2 = p.z;
            }
            const double d0 = cc0 - c0, d1 = cc1 - c1, d2 = cc2 - c2;
            const double Mc = mc.x;
            const double *qc = quad + 8 * static_cast<size_t>(c);
            q0 += qc[0] + Mc * d0 * d0;
            q1 += qc[1] + Mc * d1 * d1;
            q2 += qc[2] + Mc * d2 * d2;
            q3 += qc[3] + Mc * d0 * d1;
            q4 += qc[4] + Mc * d0 * d2;
            q5 += qc[5] + Mc * d1 * d2;
        }
    }
    double *qn = quad + 8 * static_cast<size_t>(node);
    qn[0] = q0;
    qn[1] = q1;
    qn[2] = q2;
    qn[3] = q3;
    qn[4] = q4;
    qn[5] = q5;
    qn[6] = 0.0;
    qn[7] = 0.0;
    // the walk's FP32 copy: the six moments rounded to FP32 and 1.5 tr(Q)
    // formed from them, as the quadrupole term uses them
    const float fxx = static_cast<float>(q0), fyy = static_cast<float>(q1),
                fzz = static_cast<float>(q2);
    qf[0] = make_float4(fxx, fyy, fzz, static_cast<float>(q3));
    qf[1] = make_float4(static_cast<float>(q4), static_cast<float>(q5), 1.5f * (fxx + fyy + fzz),
                        0.0f);
    if (!OCT) return;
    // third central moments T_abc = sum m s_a s_b s_c (xxx yyy zzz xxy xxz
    // xyy yyz xzz yzz xyz): leaves from their particles, internal nodes from
    // their children shifted to this centre of mass,
    //   T_abc += T^c_abc + Q^c_ab d_c + Q^c_ac d_b + Q^c_bc d_a + M_c d_a d_b d_c
    double T[10];
#pragma unroll
    for (int k = 0; k < 10; k++) T[k] = 0.0;
    if (nd.nchild == 0) {
        for (int s = nd.first; s < nd.first + nd.count; s++) {
            const double4 p = P[s];
            const double d0 = p.x - c0, d1 = p.y - c1, d2 = p.z - c2;
            const double mp = p.w;
            T[0] += mp * d0 * d0 * d0;
            T[1] += mp * d1 * d1 * d1;
            T[2] += mp * d2 * d2 * d2;
            T[3] += mp * d0 * d0 * d1;
            T[4] += mp * d0 * d0 * d2;
            T[5] += mp * d0 * d1 * d1;
            T[6] += mp * d1 * d1 * d2;
            T[7] += mp * d0 * d2 * d2;
            T[8] += mp * d1 * d2 * d2;
            T[9] += mp * d0 * d1 * d2;
        }
    } else {
        for (int c = nd.child0; c < nd.child0 + nd.nchild; c++) {
            const NodeRec &ch = nodes[c];
            const double d0 = ch.cx - c0, d1 = ch.cy - c1, d2 = ch.cz - c2;
            const double Mc = mom[c].x;
            const double *q = quad + 8 * static_cast<size_t>(c);
            const double *tc = oct + 12 * static_cast<size_t>(c);
            T[0] += tc[0] + 3.0 * q[0] * d0 + Mc * d0 * d0 * d0;
            T[1] += tc[1] + 3.0 * q[1] * d1 + Mc * d1 * d1 * d1;
            T[2] += tc[2] + 3.0 * q[2] * d2 + Mc * d2 * d2 * d2;
            T[3] += tc[3] + (q[0] * d1 + 2.0 * q[3] * d0) + Mc * d0 * d0 * d1;
            T[4] += tc[4] + (q[0] * d2 + 2.0 * q[4] * d0) + Mc * d0 * d0 * d2;
            T[5] += tc[5] + (2.0 * q[3] * d1 + q[1] * d0) + Mc * d0 * d1 * d1;
            T[6] += tc[6] + (q[1] * d2 + 2.0 * q[5] * d1) + Mc * d1 * d1 * d2;
            T[7] += tc[7] + (2.0 * q[4] * d2 + q[2] * d0) + Mc * d0 * d2 * d2;
            T[8] += tc[8] + (2.0 * q[5] * d2 + q[2] * d1) + Mc * d1 * d2 * d2;
            T[9] += tc[9] + (q[3] * d2 + q[4] * d1 + q[5] * d0) + Mc * d0 * d1 * d2;
        }
    }
    double *tn = oct + 12 * static_cast<size_t>(node);
#pragma unroll
    for (int k = 0; k < 10; k++) tn[k] = T[k];
    tn[10] = 0.0;
    tn[11] = 0.0;
    float4 *tf = reinterpret_cast<float4 *>(octf + 12 * static_cast<size_t>(node));
    tf[0] = make_float4(static_cast<float>(T[0]), static_cast<float>(T[1]), static_cast<float>(T[2]),
                        static_cast<float>(T[3]));
    tf[1] = make_float4(static_cast<float>(T[4]), static_cast<float>(T[5]), static_cast<float>(T[6]),
                        static_cast<float>(T[7]));
    tf[2] = make_float4(static_cast<float>(T[8]), static_cast<float>(T[9]), 0.0f, 0.0f);
}

// ------------------------------------------------------------------ build, all levels at once
// The same node set and numbering as the level-by-level build, found from
// the sorted keys alone.  L_i = number of leading octal digits particle i
// shares with particle i-1 (L_0 = L_n = -1): particle i starts a cell at
// every depth d > L_i.  The cell of depth d containing i holds at most
// leaf_cap particles exactly when some a <= i < b with b - a <= leaf_cap has
// L_a < d and L_b < d, so the first such depth is
//   D_i = min over those (a, b) of max(L_a, L_b) + 1     (clamped to 1..21:
// depth 0 is always subdivided, depth 21 never).  The nodes are the pairs
// (i, d) with L_i < d <= D_i: level d in order of first particle i, internal
// when d < D_i, its first child being (i, d + 1).
constexpr int FB_BLOCK = 256;
constexpr int FB_PER = 1;                   // particles per thread
constexpr int FB_TILE = FB_BLOCK * FB_PER;  // particles per workgroup
constexpr int FB_DEPTHS = MAXL + 1;
constexpr int FB_WAVES = FB_BLOCK / 64;
constexpr int FB_MAX_LEAF = 16;             // window search cost grows as leaf_cap^2
constexpr int FB_SCAN = 64;                 // L entries a thread scans before leaving a cell open
constexpr int FB_PAD = 4;                   // -1 entries after L_n (the scan reads 4 at a time)

__device__ __forceinline__ int shared_digits(long long a, long long b)
{
    const unsigned long long x = static_cast<unsigned long long>(a ^ b);
    if (x == 0) return MAXL;
    const int h = 63 - __clzll(x);  // highest differing bit (keys use bits 0..62)
    return MAXL - 1 - h / 3;
}

__global__ void k_fb_digits(const long long *__restrict__ keys, int n, signed char *__restrict__ L)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n + FB_PAD) return;
    L[i] = static_cast<signed char>((i == 0 || i >= n) ? -1 : shared_digits(keys[i - 1], keys[i]));
}

__global__ void k_fb_leafdepth(const signed char *__restrict__ L, int n, int leaf_cap,
                               signed char *__restrict__ D)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int best = MAXL + 1;
    const int a0 = i - leaf_cap + 1 > 0 ? i - leaf_cap + 1 : 0;
    for (int a = a0; a <= i; a++) {
        const int la = L[a];
        const int bmax = a + leaf_cap < n ? a + leaf_cap : n;
        for (int b = i + 1; b <= bmax; b++) {
            const int lb = L[b];
            const int v = (la > lb ? la : lb) + 1;
            best = v < best ? v : best;
        }
    }
    D[i] = static_cast<signed char>(best < 1 ? 1 : (best > MAXL ? MAXL : best));
}

__device__ __forceinline__ bool fb_node(int l, int dd, int d) { return l < d && d <= dd; }

// nodes per (depth, tile), depth-major for the scan
__global__ __launch_bounds__(FB_BLOCK) void k_fb_count(const signed char *__restrict__ L,
                                                       const signed char *__restrict__ D, int n,
                                                       int ntiles, unsigned *__restrict__ cnt)
{
    __shared__ unsigned c[FB_DEPTHS];
    if (threadIdx.x < FB_DEPTHS) c[threadIdx.x] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    for (int r = 0; r < FB_PER; r++) {
        const int i = blockIdx.x * FB_TILE + r * FB_BLOCK + threadIdx.x;
        const int l = i < n ? L[i] : MAXL + 1;
        const int dd = i < n ? D[i] : -1;
        for (int d = 0; d < FB_DEPTHS; d++) {
            const unsigned long long m = __ballot(fb_node(l, dd, d));
            if (lane == 0 && m) atomicAdd(&c[d], static_cast<unsigned>(__popcll(m)));
        }
    }
    __syncthreads();
    if (threadIdx.x < FB_DEPTHS) cnt[threadIdx.x * ntiles + blockIdx.x] = c[threadIdx.x];
}

// level offsets (first node of each depth) and the node count
__global__ void k_fb_levels(const unsigned *__restrict__ scanned, int ntiles,
                            const unsigned *__restrict__ total, unsigned *__restrict__ out)
{
    const int d = threadIdx.x;
    if (d < FB_DEPTHS) out[d] = scanned[d * ntiles];
    if (d == FB_DEPTHS) out[d] = *total;
}

// first position after i whose depth-d prefix differs from i's (galloping)
__device__ int cell_end(const long long *__restrict__ keys, int n, int i, int shift)
{
    const long long pre = keys[i] >> shift;
    int lo = i + 1, hi = n, step = 1;  // keys[j] share the prefix for i <= j < lo
    for (;;) {
        const int j = i + step;
        if (j >= n) break;
        if ((keys[j] >> shift) != pre) {
            hi = j;
            break;
        }
        lo = j + 1;
        step <<= 1;
    }
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((keys[mid] >> shift) == pre) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(FB_BLOCK) void k_fb_emit(const long long *__restrict__ keys,
                                                      const signed char *__restrict__ L,
                                                      const signed char *__restrict__ D, int n,
                                                      int ntiles,
                                                      const unsigned *__restrict__ scanned,
                                                      double w2, NodeRec *__restrict__ nodes,
                                                      unsigned char *__restrict__ fix)
{
    __shared__ unsigned base[FB_DEPTHS][FB_PER][FB_WAVES];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const unsigned long long lt = (1ull << lane) - 1ull;
    int l[FB_PER], dd[FB_PER];
    for (int r = 0; r < FB_PER; r++) {
        const int i = blockIdx.x * FB_TILE + r * FB_BLOCK + threadIdx.x;
        l[r] = i < n ? L[i] : MAXL + 1;
        dd[r] = i < n ? D[i] : -1;
        for (int d = 0; d < FB_DEPTHS; d++) {
            const unsigned long long m = __ballot(fb_node(l[r], dd[r], d));
            if (lane == 0) base[d][r][wave] = static_cast<unsigned>(__popcll(m));
        }
    }
    __syncthreads();
    // exclusive offsets within the tile, in particle order (r, wave, lane)
    if (threadIdx.x < FB_DEPTHS) {
        const int d = threadIdx.x;
        unsigned acc = scanned[d * ntiles + blockIdx.x];
        for (int r = 0; r < FB_PER; r++)
            for (int w = 0; w < FB_WAVES; w++) {
                const unsigned c = base[d][r][w];
                base[d][r][w] = acc;
                acc += c;
            }
    }
    __syncthreads();
    for (int r = 0; r < FB_PER; r++) {
        const int i = blockIdx.x * FB_TILE + r * FB_BLOCK + threadIdx.x;
        int below = -1;  // index of node (i, d + 1)
        // Cell ends and child counts from a forward scan of L: the cell of
        // depth d starting at i ends at the first j > i with L_j < d, and its
        // children start at i and at the positions before that end with
        // L_j == d (all positions inside have L_j >= d).  Depths are visited
        // deepest first, so the scan resumes where the deeper cell ended; past
        // FB_SCAN positions the cell is left to k_fb_fix.
        int j = i + 1;
        bool scan = true;
        for (int d = FB_DEPTHS - 1; d >= 0; d--) {
            const bool is = fb_node(l[r], dd[r], d);
            const unsigned long long m = __ballot(is);
            if (!is) continue;
            const int idx = static_cast<int>(base[d][r][wave] + __popcll(m & lt));
            const int shift = 3 * (MAXL - d);
            int end = -1, kids = 1;
            if (d > 0 && scan) {
                while (j - i <= FB_SCAN) {
                    const int v0 = L[j], v1 = L[j + 1], v2 = L[j + 2], v3 = L[j + 3];
                    if (v0 < d) { end = j; break; }
                    kids += v0 == d ? 1 : 0;
                    if (v1 < d) { end = j + 1; break; }
                    kids += v1 == d ? 1 : 0;
                    if (v2 < d) { end = j + 2; break; }
                    kids += v2 == d ? 1 : 0;
                    if (v3 < d) { end = j + 3; break; }
                    kids += v3 == d ? 1 : 0;
                    j += 4;
                }
                if (end >= 0) j = end;
                else scan = false;
            }
            NodeRec nd;
            nd.cx = nd.cy = nd.cz = 0.0;
            // (2w / (2 << d))^2 == w^2 * 4^-d exactly (power-of-two scaling)
            nd.l2 = __builtin_amdgcn_ldexp(w2, -2 * d);
            nd.prefix = keys[i] >> shift;
            nd.depth = d;
            nd.first = i;
            // a cell the scan did not close gets its count and child count
            // from searches on the keys in k_fb_fix, so that the few threads
            // with long searches do not hold their workgroups here
            nd.count = d == 0 ? n : (end >= 0 ? end - i : 0);
            nd.fmass = 0.0f;
            fix[idx] = end < 0 ? 1 : 0;
            if (d < dd[r]) {
                nd.child0 = below;
                nd.nchild = end < 0 ? 0 : kids;
            } else {
                nd.child0 = -1;
                nd.nchild = 0;
            }
            nodes[idx] = nd;
            below = idx;
        }
    }
}

// First position after i whose depth-d prefix (key >> shift) differs from
// i's, or n: eight independent probes per step, first at i + 8^k (k = 0..7),
// then splitting the bracket nine ways, so that a cell of C particles costs
// about 1 + log9 C dependent loads instead of the 2 log2 C of a galloping
// binary search.  The prefix is constant on [i, end) (sorted keys).
__device__ int cell_end8(const long long *__restrict__ keys, int n, int i, int shift)
{
    const long long pre = keys[i] >> shift;
    int lo = i + 1, hi = n;  // invariant: the end lies in [lo, hi]
    {
        bool same[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const long long q = static_cast<long long>(i) + (1ll << (3 * k));
            same[k] = q < n && (keys[q] >> shift) == pre;
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const long long q = static_cast<long long>(i) + (1ll << (3 * k));
            if (same[k]) lo = static_cast<int>(q) + 1;
            else if (q < hi) hi = static_cast<int>(q);
        }
        if (lo > hi) lo = hi;  // (cannot happen for sorted keys)
    }
    while (lo < hi) {
        const long long len = hi - lo;
        int q[8];
        bool same[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            q[k] = lo + static_cast<int>(len * (k + 1) / 9);  // in [lo, hi)
            same[k] = (keys[q[k]] >> shift) == pre;
        }
        int nlo = lo, nhi = hi;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            if (same[k]) nlo = q[k] + 1 > nlo ? q[k] + 1 : nlo;
            else nhi = q[k] < nhi ? q[k] : nhi;
        }
        lo = nlo < nhi ? nlo : nhi;
        hi = nhi;
    }
    return lo;
}

// The cells k_fb_emit left open (fix[node] != 0: the root, and cells whose
// end lies more than FB_SCAN positions on): the count from the end of the
// cell's key prefix; for an internal node the child count from the next
// level's records: its children are the depth-(d + 1) nodes whose first
// particle lies in the cell, contiguous from child0 (at most 8) and ending
// with the level (lev[d + 2]: first node of depth d + 2, or the node count).
__global__ void k_fb_fix(NodeRec *__restrict__ nodes, int nn, const long long *__restrict__ keys,
                         int n, const unsigned char *__restrict__ fix,
                         const unsigned *__restrict__ lev)
{
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= nn || !fix[node]) return;
    NodeRec &nd = nodes[node];
    const int i = nd.first, d = nd.depth;
    const int end = d == 0 ? n : cell_end8(keys, n, i, 3 * (MAXL - d));
    nd.count = end - i;
    if (nd.child0 >= 0) {
        const int c0 = nd.child0;
        const int le = static_cast<int>(lev[d + 2]);
        int first[8];
#pragma unroll
        for (int k = 0; k < 8; k++) first[k] = c0 + k < le ? nodes[c0 + k].first : end;
        int nc = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) nc += first[k] < end ? 1 : 0;
        nd.nchild = nc;
    }
}

// ------------------------------------------------------------------ walk
constexpr int WK_BLOCK = 256;
constexpr int WK_WAVES = WK_BLOCK / 64;
constexpr int WK_STACK = 192;  // >= 7 pending siblings x 21 levels + 8

struct WalkArgs {
    double theta2;
    double eps2;
    double G;
    double phi2;     // group walk: squared expansion ratio (0: no local expansions)
    double thetaf2;  // squared opening angle for far nodes (expanded ones)
};

__device__ __forceinline__ void pp_force(double xi0, double xi1, double xi2, double px, double py,
                                         double pz, double pm, double eps2, double &a0,
                                         double &a1, double &a2)
{
    const double R0 = xi0 - px, R1 = xi1 - py, R2 = xi2 - pz;
    const double r2 = R0 * R0 + R1 * R1 + R2 * R2 + eps2;
    const float ri = __builtin_amdgcn_rsqf(static_cast<float>(r2));
    const double w = static_cast<double>(static_cast<float>(pm) * (ri * ri * ri));
    a0 = __builtin_fma(-w, R0, a0);
    a1 = __builtin_fma(-w, R1, a1);
    a2 = __builtin_fma(-w, R2, a2);
}

// Quadrupole term of an accepted node (R = x_i - com, r2 = |R|^2 + eps^2),
// without the factor G: 3 Q R / r^5 - (15/2) (R.Q.R) R / r^7 + (3/2) tr(Q) R / r^5.
__device__ __forceinline__ void quad_term(const float *__restrict__ q, double R0, double R1,
                                          double R2, float ri, float &o0, float &o1, float &o2)
{
    const float r0 = static_cast<float>(R0), r1 = static_cast<float>(R1),
                r2 = static_cast<float>(R2);
    // (q: the moments in FP32 and 1.5 tr(Q), written by k_multipole)
    const float qxx = q[0], qyy = q[1], qzz = q[2], qxy = q[3], qxz = q[4], qyz = q[5];
    const float QR0 = __builtin_fmaf(qxz, r2, __builtin_fmaf(qxy, r1, qxx * r0));
    const float QR1 = __builtin_fmaf(qyz, r2, __builtin_fmaf(qyy, r1, qxy * r0));
    const float QR2 = __builtin_fmaf(qzz, r2, __builtin_fmaf(qyz, r1, qxz * r0));
    const float RQR = __builtin_fmaf(r2, QR2, __builtin_fmaf(r1, QR1, r0 * QR0));
    const float ir2 = ri * ri;
    const float ir5 = ir2 * ir2 * ri;
    const float sc = ir5 * __builtin_fmaf(-7.5f * RQR, ir2, q[6]);
    const float k3 = 3.0f * ir5;
    o0 = __builtin_fmaf(k3, QR0, sc * r0);
    o1 = __builtin_fmaf(k3, QR1, sc * r1);
    o2 = __builtin_fmaf(k3, QR2, sc * r2);
}

// Octupole term of an accepted node (R = x_i - com, r2 = |R|^2 + eps^2),
// without the factor G, from its third central moments T (FP32, xxx yyy zzz
// xxy xxz xyy yyz xzz yzz xyz):
//   -(35/2) u R / r^9 + (15/2) ((t.R) R + w) / r^7 - (3/2) t / r^5,
// w_d = T_abd R_a R_b, u = w.R, t_d = T_aad.
__device__ __forceinline__ void oct_term(const float *__restrict__ T, double R0, double R1,
                                         double R2, float ri, float &o0, float &o1, float &o2)
{
    const float r0 = static_cast<float>(R0), r1 = static_cast<float>(R1),
                r2 = static_cast<float>(R2);
    const float Txxx = T[0], Tyyy = T[1], Tzzz = T[2], Txxy = T[3], Txxz = T[4], Txyy = T[5],
                Tyyz = T[6], Txzz = T[7], Tyzz = T[8], Txyz = T[9];
    const float r00 = r0 * r0, r11 = r1 * r1, r22 = r2 * r2;
    const float r01 = 2.0f * r0 * r1, r02 = 2.0f * r0 * r2, r12 = 2.0f * r1 * r2;
    const float w0 = Txxx * r00 + Txyy * r11 + Txzz * r22 + Txxy * r01 + Txxz * r02 + Txyz * r12;
    const float w1 = Txxy * r00 + Tyyy * r11 + Tyzz * r22 + Txyy * r01 + Txyz * r02 + Tyyz * r12;
    const float w2 = Txxz * r00 + Tyyz * r11 + Tzzz * r22 + Txyz * r01 + Txzz * r02 + Tyzz * r12;
    const float t0 = Txxx + Txyy + Txzz, t1 = Txxy + Tyyy + Tyzz, t2 = Txxz + Tyyz + Tzzz;
    const float u = w0 * r0 + w1 * r1 + w2 * r2;
    const float tR = t0 * r0 + t1 * r1 + t2 * r2;
    const float ir2 = ri * ri;
    const float ir5 = ir2 * ir2 * ri;
    const float ir7 = ir5 * ir2;
    const float su = (-17.5f * u * ir2 + 7.5f * tR) * ir7;
    o0 = su * r0 + 7.5f * w0 * ir7 - 1.5f * t0 * ir5;
    o1 = su * r1 + 7.5f * w1 * ir7 - 1.5f * t1 * ir5;
    o2 = su * r2 + 7.5f * w2 * ir7 - 1.5f * t2 * ir5;
}

// QUAD: add the quadrupole term of accepted nodes.  FEAT: do not add it but
// accumulate per-target features of the walk for the learned correction,
// 16 FP32 per particle in original order, without the factor G:
// (quadrupole sum x3, sum (M l^3 / d^6) R x3, sum M l^2 / d^4,
// sum M l^3 / d^5, accepted nodes, particle interactions, octupole sum x3,
// 0, 0, 0) with l the node's side, R = x_i - com and d = |R| (softened).
template <bool QUAD, bool FEAT>
__global__ __launch_bounds__(WK_BLOCK) void k_walk(const NodeRec *__restrict__ nodes,
                                                   const double4 *__restrict__ P,
                                                   const long long *__restrict__ skeys,
                                                   const int *__restrict__ sidx, int n,
                                                   WalkArgs args, double *__restrict__ a,
                                                   unsigned long long *__restrict__ n_int,
                                                   const float *__restrict__ quad,
                                                   const float *__restrict__ oct,
                                                   float *__restrict__ feat)
{
    __shared__ int stk_node[WK_WAVES][WK_STACK];
    __shared__ unsigned long long stk_mask[WK_WAVES][WK_STACK];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int s = (blockIdx.x * WK_WAVES + wave) * 64 + lane;
    const bool valid = s < n;
    const unsigned long long all = __ballot(valid);
    if (all == 0) return;

    double xi0 = 0.0, xi1 = 0.0, xi2 = 0.0;
    long long key = 0;
    if (valid) {
        const double4 p = P[s];
        xi0 = p.x;
        xi1 = p.y;
        xi2 = p.z;
        key = skeys[s];
    }
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    unsigned cnt = 0;             // per-lane interactions of the rare opened-leaf path
    unsigned long long wcnt = 0;  // the wavefront's other interactions (uniform)
    float fq0 = 0.0f, fq1 = 0.0f, fq2 = 0.0f, fs2 = 0.0f, fs3 = 0.0f;
    float fv0 = 0.0f, fv1 = 0.0f, fv2 = 0.0f;
    float fo0 = 0.0f, fo1 = 0.0f, fo2 = 0.0f;
    unsigned nacc = 0, npp = 0;
    int *snode = stk_node[wave];
    unsigned long long *smask = stk_mask[wave];

    int sp = 0;
    {
        const int c0 = nodes[0].child0, nc = nodes[0].nchild;
        for (int c = nc - 1; c >= 0; c--) {
            if (lane == 0) {
                snode[sp] = c0 + c;
                smask[sp] = all;
            }
            sp++;
        }
    }
    __builtin_amdgcn_wave_barrier();
    while (sp > 0) {
        sp--;
        const int node = __builtin_amdgcn_readfirstlane(snode[sp]);
        const unsigned long long M =
            (static_cast<unsigned long long>(__builtin_amdgcn_readfirstlane(
                 static_cast<int>(smask[sp] >> 32)))
             << 32) |
            static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(smask[sp])));
        __builtin_amdgcn_wave_barrier();
        const NodeRec nd = nodes[node];
        // (M is uniform: the lane's bit as a per-lane predicate, no VALU work)
        const bool act = __builtin_amdgcn_inverse_ballot_w64(M);
        unsigned long long O = 0;  // lanes that open this node
        // The VALU executes an instruction for the whole wavefront whatever
        // its lane mask, so the terms are formed branch-free for every lane
        // and lanes that do not take them add an exact zero (the running sums
        // start at +0 and a sum of products that cancels is +0, never -0, so
        // adding fma(-0, R, a) = a leaves every bit unchanged).
        if (nd.nchild == 0 && nd.count == 1) {
            // single-particle leaf: its monopole is the particle itself
            const bool use = act && nd.first != s;
            wcnt += static_cast<unsigned>(__popcll(__ballot(use)));
            {
                const double4 p = P[nd.first];
                const double R0 = xi0 - p.x, R1 = xi1 - p.y, R2 = xi2 - p.z;
                const double r2 = R0 * R0 + R1 * R1 + R2 * R2 + args.eps2;
                const float ri = __builtin_amdgcn_rsqf(static_cast<float>(r2));
                const float mr3 = static_cast<float>(p.w) * (ri * ri * ri);
                const double w = static_cast<double>(use ? mr3 : 0.0f);
                a0 = __builtin_fma(-w, R0, a0);
                a1 = __builtin_fma(-w, R1, a1);
                a2 = __builtin_fma(-w, R2, a2);
                if (FEAT) npp += use ? 1u : 0u;
            }
            continue;
        }
        const bool included = (key >> (3 * (MAXL - nd.depth))) == nd.prefix;
        const double R0 = xi0 - nd.cx, R1 = xi1 - nd.cy, R2 = xi2 - nd.cz;
        const double d2 = R0 * R0 + R1 * R1 + R2 * R2;
        const double l2 = nd.l2;
        const bool accept = act && !included && l2 < args.theta2 * d2;
        const unsigned long long AM = __ballot(accept);
        wcnt += static_cast<unsigned>(__popcll(AM));
        O = M & ~AM;  // act && !accept
        // (no branch around the terms: a visit where no lane accepts is rare
        // enough that computing them anyway is cheaper than merging the sums)
        {
            const float ri = __builtin_amdgcn_rsqf(static_cast<float>(d2 + args.eps2));
            const float mr3 = nd.fmass * (ri * ri * ri);
            const double w = static_cast<double>(accept ? mr3 : 0.0f);
            a0 = __builtin_fma(-w, R0, a0);
            a1 = __builtin_fma(-w, R1, a1);
            a2 = __builtin_fma(-w, R2, a2);
            if (QUAD || FEAT) {
                float o0, o1, o2;
                quad_term(quad + 8 * static_cast<size_t>(node), R0, R1, R2, ri, o0, o1, o2);
                if (QUAD) {
                    a0 += static_cast<double>(accept ? o0 : 0.0f);
                    a1 += static_cast<double>(accept ? o1 : 0.0f);
                    a2 += static_cast<double>(accept ? o2 : 0.0f);
                } else {
                    // M l^2 / d^4 and M l^3 / d^5 (l: cell side, d: distance to the COM)
                    const float lr = static_cast<float>(__builtin_sqrt(l2)) * ri;
                    const float mr2 = mr3 * static_cast<float>(__builtin_sqrt(d2 + args.eps2));
                    const float lr3 = lr * lr * lr;
                    const float k3 = mr3 * lr3;
                    float t0, t1, t2;
                    oct_term(oct + 12 * static_cast<size_t>(node), R0, R1, R2, ri, t0, t1, t2);
                    if (accept) {
                        fq0 += o0;
                        fq1 += o1;
                        fq2 += o2;
                        fo0 += t0;
                        fo1 += t1;
                        fo2 += t2;
                        fs2 = __builtin_fmaf(mr2, lr * lr, fs2);
                        fs3 = __builtin_fmaf(mr2, lr3, fs3);
                        fv0 = __builtin_fmaf(k3, static_cast<float>(R0), fv0);
                        fv1 = __builtin_fmaf(k3, static_cast<float>(R1), fv1);
                        fv2 = __builtin_fmaf(k3, static_cast<float>(R2), fv2);
                        nacc++;
                    }
                }
            }
        }
        if (O == 0) continue;
        if (nd.nchild == 0) {
            const bool mine = __builtin_amdgcn_inverse_ballot_w64(O);
            for (int t = nd.first; t < nd.first + nd.count; t++) {
                const double4 p = P[t];
                if (mine && t != s) {
                    pp_force(xi0, xi1, xi2, p.x, p.y, p.z, p.w, args.eps2, a0, a1, a2);
                    cnt++;
                    npp++;
                }
            }
        } else if (sp + nd.nchild <= WK_STACK) {
            // children pushed last-to-first (child0 on top), one lane each
            if (lane < nd.nchild) {
                snode[sp + lane] = nd.child0 + nd.nchild - 1 - lane;
                smask[sp + lane] = O;
            }
            sp += nd.nchild;
            __builtin_amdgcn_wave_barrier();
        } else if (lane == 0 && n_int) {
            // cannot happen for depth <= 21 (at most 7 pending siblings per
            // level); flagged through the counter instead of overflowing
            atomicAdd(n_int, 1ull << 62);
        }
    }
    if (valid) {
        const int i = sidx[s];
        a[3 * i + 0] = args.G * a0;
        a[3 * i + 1] = args.G * a1;
        a[3 * i + 2] = args.G * a2;
        if (FEAT) {
            float4 *f = reinterpret_cast<float4 *>(feat + 16 * static_cast<size_t>(i));
            f[0] = make_float4(fq0, fq1, fq2, fv0);
            f[1] = make_float4(fv1, fv2, fs2, fs3);
            f[2] = make_float4(static_cast<float>(nacc), static_cast<float>(npp), fo0, fo1);
            f[3] = make_float4(fo2, 0.0f, 0.0f, 0.0f);
        }
    }
    // interaction count: one atomic per wave
    unsigned long long tot = cnt;
    for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
    if (lane == 0 && n_int) atomicAdd(n_int, tot + wcnt);
}

// ------------------------------------------------------------------ group walk
// One wavefront walks the tree for a group of at most 64 consecutive
// particles in sorted order (formed from the tree by k_group_runs, below): a
// node is accepted for the whole group when it holds none
// of the group's particles and l^2 < theta^2 d^2, d the distance from its
// centre of mass to the group's bounding box -- a test every particle of the
// group would pass on its own (d_i >= d), so the interaction set refines the
// per-particle walk's.  The traversal is breadth-parallel: each pass takes up
// to 64 pending nodes from the wave's LDS stack, one per lane; accepted nodes
// and single-particle leaves go to LDS interaction lists, opened nodes push
// their children.  Each full list (64 entries) is then summed by every lane
// for its own particle like a direct-sum tile: coordinates relative to the
// group centre in FP32, 64-term FP32 partial sums added to FP64 totals.
// The pending frontier of a group reaches 550 (median) to ~900 nodes at
// N = 2^20 - 2^21.  A 1024-entry stack (4 KB; 9.5 KB of LDS per wavefront)
// keeps most passes full and ~16 wavefronts resident per CU: at N = 2^21,
// theta 0.5, the launch takes 5.2 ms against 6.0 ms with 2048 entries (fewer
// resident wavefronts) and 5.8 / 12.2 ms with 768 / 512 (more single-node
// passes once the frontier reaches the batch limit below).  One wavefront per
// workgroup: 6.0 ms against 6.7 with two and 8.2 with four (2048 entries).
constexpr int GW_WAVES = 1;
constexpr int GW_BLOCK = 64 * GW_WAVES;
constexpr int GW_STACK = 1024;
constexpr int GW_LIST = 128;
// batch passes keep the stack below this; single-node passes then grow it by
// at most 7 per level of the remaining depth
constexpr int GW_BATCH_LIMIT = GW_STACK - 8 * (MAXL + 2);

// Wave-wide minimum / maximum of a double, every lane active: within each
// row of 16 lanes by DPP moves (quad swaps, then row rotations by 4 and 8),
// then the four rows' results read from lanes 0, 16, 32, 48.  No LDS
// traffic; min and max are exact, so the order does not matter.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = static_cast<int>(b), hi = static_cast<int>(b >> 32);
    const int l = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
    const int h = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
    return __longlong_as_double((static_cast<long long>(h) << 32) | static_cast<unsigned>(l));
}
__device__ __forceinline__ double lane_f64(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), lane);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
constexpr int DPP_QUAD_XOR1 = 0xB1;  // quad_perm [1, 0, 3, 2]
constexpr int DPP_QUAD_XOR2 = 0x4E;  // quad_perm [2, 3, 0, 1]
constexpr int DPP_ROW_ROR4 = 0x124;
constexpr int DPP_ROW_ROR8 = 0x128;
__device__ __forceinline__ double wave_min(double v)
{
    v = fmin(v, dpp_f64<DPP_QUAD_XOR1>(v));
    v = fmin(v, dpp_f64<DPP_QUAD_XOR2>(v));
    v = fmin(v, dpp_f64<DPP_ROW_ROR4>(v));
    v = fmin(v, dpp_f64<DPP_ROW_ROR8>(v));
    return fmin(fmin(lane_f64(v, 0), lane_f64(v, 16)), fmin(lane_f64(v, 32), lane_f64(v, 48)));
}
__device__ __forceinline__ double wave_max(double v)
{
    v = fmax(v, dpp_f64<DPP_QUAD_XOR1>(v));
    v = fmax(v, dpp_f64<DPP_QUAD_XOR2>(v));
    v = fmax(v, dpp_f64<DPP_ROW_ROR4>(v));
    v = fmax(v, dpp_f64<DPP_ROW_ROR8>(v));
    return fmax(fmax(lane_f64(v, 0), lane_f64(v, 16)), fmax(lane_f64(v, 32), lane_f64(v, 48)));
}
// A value every lane holds alike, moved to scalar registers (the compiler
// cannot see that a cross-lane reduction is uniform).
__device__ __forceinline__ double uniform(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane(static_cast<int>(b));
    const int hi = __builtin_amdgcn_readfirstlane(static_cast<int>(b >> 32));
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// Quadrupole term in FP32, R = x_i - com (the same expression as quad_term).
__device__ __forceinline__ void quad_termf(const float *__restrict__ q, float r0, float r1,
                                           float r2, float ri, float &o0, float &o1, float &o2)
{
    const float qxx = q[0], qyy = q[1], qzz = q[2], qxy = q[3], qxz = q[4], qyz = q[5];
    const float QR0 = __builtin_fmaf(qxz, r2, __builtin_fmaf(qxy, r1, qxx * r0));
    const float QR1 = __builtin_fmaf(qyz, r2, __builtin_fmaf(qyy, r1, qxy * r0));
    const float QR2 = __builtin_fmaf(qzz, r2, __builtin_fmaf(qyz, r1, qxz * r0));
    const float RQR = __builtin_fmaf(r2, QR2, __builtin_fmaf(r1, QR1, r0 * QR0));
    const float ir2 = ri * ri;
    const float ir5 = ir2 * ir2 * ri;
    const float sc = ir5 * __builtin_fmaf(-7.5f * RQR, ir2, q[6]);
    const float k3 = 3.0f * ir5;
    o0 = __builtin_fmaf(k3, QR0, sc * r0);
    o1 = __builtin_fmaf(k3, QR1, sc * r1);
    o2 = __builtin_fmaf(k3, QR2, sc * r2);
}

// Walk groups from the tree: the children with at most 64 particles of each
// node with more are packed, in order, into runs of at most 64 contiguous
// particles (a root of n <= 64 particles is one group; a leaf of more than
// 64 particles gives runs of 64).  A group's box can cover cells outside the
// group, whose nodes the whole group must then open down to the leaves: a
// run whose box is wider than `cap` in some axis (a few times the width of a
// 64-particle group at the mean density) gives one group per particle --
// sparse outliers, for which that is the per-particle walk.
// The nodes that make groups: the root, and every node of more than 64
// particles (flagged by the moments pass; scanned; then the list at the
// scanned offsets, a node being flagged where its offset and the next differ).
__global__ void k_group_parent_list(int num_nodes, const unsigned *__restrict__ off,
                                    const unsigned *__restrict__ total, int *__restrict__ parents)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= num_nodes) return;
    const unsigned o = off[p], o1 = p + 1 < num_nodes ? off[p + 1] : *total;
    if (o1 > o) parents[o] = p;
}

// One wavefront per parent node: its runs are formed in the same order by
// every lane; each run's particles are loaded one per lane and its box is
// reduced across the wave.  Groups are counted in a first launch (WRITE =
// false: count[i] := their number, wide[i] := bit r set when run r of an
// internal node is wider than the cap) and written in a second from the
// exclusive scan of the counts (WRITE = true: groups[off[i] ...]), which
// takes the internal nodes' widths from wide[i].
__device__ __forceinline__ bool run_is_wide(double4 q, double cap)
{
    // q: particle f + lane of the run (its first for lanes past the run)
    const double w0 = wave_max(q.x) - wave_min(q.x);
    const double w1 = wave_max(q.y) - wave_min(q.y);
    const double w2 = wave_max(q.z) - wave_min(q.z);
    return w0 > cap || w1 > cap || w2 > cap;
}

template <bool WRITE>
__device__ __forceinline__ void emit_group(bool wide, int f, int c, int lane,
                                           int2 *__restrict__ groups, unsigned &at)
{
    if (wide) {
        if (WRITE && lane < c) groups[at + lane] = make_int2(f + lane, 1);
        at += c;
    } else {
        if (WRITE && lane == 0) groups[at] = make_int2(f, c);
        at += 1;
    }
}

template <bool WRITE>
__global__ __launch_bounds__(256) void k_group_runs(const NodeRec *__restrict__ nodes,
                                                    const int *__restrict__ parents, int num_parents,
                                                    int n, const double4 *__restrict__ P, double cap,
                                                    unsigned *__restrict__ count_or_off,
                                                    unsigned *__restrict__ wide,
                                                    int2 *__restrict__ groups)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (i >= num_parents) return;
    const int p = parents[i];
    unsigned at = WRITE ? count_or_off[i] : 0u;
    const NodeRec nd = nodes[p];
    if (nd.count <= 64) {
        // (the root of at most 64 particles: one group)
        if (WRITE && lane == 0) groups[at] = make_int2(0, n);
        at += 1;
    } else if (nd.nchild == 0) {
        // a leaf of more than 64 particles: consecutive runs of 64
        for (int f = nd.first; f < nd.first + nd.count; f += 64) {
            const int c = min(64, nd.first + nd.count - f);
            const bool w = run_is_wide(P[f + (lane < c ? lane : 0)], cap);
            emit_group<WRITE>(w, f, c, lane, groups, at);
        }
    } else {
        // the children's (first, count), one child per lane; the runs packed
        // in child order, run r's (first, count) kept by lane r (at most 8
        // runs: one per child)
        int cf = 0, cc = 0;
        if (lane < nd.nchild) {
            cf = nodes[nd.child0 + lane].first;
            cc = nodes[nd.child0 + lane].count;
        }
        int start = -1, cnt = 0, nr = 0, rf = 0, rn = 0;
        for (int k = 0; k < nd.nchild; k++) {
            const int kf = __shfl(cf, k, 64), kc = __shfl(cc, k, 64);
            if (kc > 64 || cnt + kc > 64) {
                if (cnt > 0) {
                    if (lane == nr) {
                        rf = start;
                        rn = cnt;
                    }
                    nr++;
                }
                start = -1;
                cnt = 0;
            }
            if (kc <= 64) {
                if (cnt == 0) start = kf;
                cnt += kc;
            }
        }
        if (cnt > 0) {
            if (lane == nr) {
                rf = start;
                rn = cnt;
            }
            nr++;
        }
        if (WRITE) {
            const unsigned mask = wide[i];
            for (int r = 0; r < nr; r++)
                emit_group<WRITE>((mask >> r) & 1u, __shfl(rf, r, 64), __shfl(rn, r, 64), lane,
                                  groups, at);
        } else {
            // every run's particles requested before any is reduced
            double4 q[8];
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const int f = __shfl(rf, r, 64), c = __shfl(rn, r, 64);
                q[r] = r < nr ? P[f + (lane < c ? lane : 0)] : make_double4(0.0, 0.0, 0.0, 0.0);
            }
            unsigned mask = 0;
#pragma unroll
            for (int r = 0; r < 8; r++)
                if (r < nr) {
                    const bool w = run_is_wide(q[r], cap);
                    mask |= w ? 1u << r : 0u;
                    emit_group<WRITE>(w, __shfl(rf, r, 64), __shfl(rn, r, 64), lane, groups, at);
                }
            if (lane == 0) wide[i] = mask;
        }
    }
    if (!WRITE && lane == 0) count_or_off[i] = at;
}

// Moments of the far nodes' fields about the group centre c, for the
// second-order local expansion.  A source of mass M at y = c - com
// (s2 = |y|^2 + eps^2, q1 = M s2^-3/2, q2 = 3 M s2^-5/2, q3 = 15 M s2^-7/2)
// gives the field a(c + d) = a0 + J d + (1/2) H[d, d] + O(|d|^3) with
//   a0_k = -q1 y_k,  J_kl = q2 y_k y_l - q1 delta_kl,
//   H_klm = q2 (delta_kl y_m + delta_km y_l + delta_lm y_k) - q3 y_k y_l y_m,
// so the sums Q = sum q1, A = sum -q1 y, S = sum q2 y, P = sum q2 y y and
// T = sum q3 y y y (symmetric: 6 and 10 components) determine it.
struct FarMoments {
    float Q, A0, A1, A2, S0, S1, S2;
    float Pxx, Pyy, Pzz, Pxy, Pxz, Pyz;
    float Txxx, Tyyy, Tzzz, Txxy, Txxz, Txyy, Tyyz, Txzz, Tyzz, Txyz;
};

// One far node: its monopole moments, and the value of its quadrupole field
// at the centre (R = c - com = y) added to a0.  The moments are formed from
// u = y / s (|u| <= 1): P = 3 q1 u u, S = 3 q1 s^-1 u, T = 15 q1 s^-1 u u u,
// so no intermediate exceeds M s^-4 (q3 itself, M s^-7, would overflow FP32
// for heavy nodes a few 1e-6 away without softening).
__device__ __forceinline__ void far_add(FarMoments &F, float y0, float y1, float y2, float m,
                                        float eps2, const float *__restrict__ quad)
{
    float s2 = __builtin_fmaf(y0, y0, eps2);
    s2 = __builtin_fmaf(y1, y1, s2);
    s2 = __builtin_fmaf(y2, y2, s2);
    const float ri = __builtin_amdgcn_rsqf(s2);
    const float u0 = y0 * ri, u1 = y1 * ri, u2 = y2 * ri;
    const float q1 = (m * ri) * (ri * ri);
    const float c1 = 3.0f * q1, c2 = c1 * ri, c3 = 5.0f * c2;
    float o0, o1, o2;
    quad_termf(quad, y0, y1, y2, ri, o0, o1, o2);
    F.Q += q1;
    F.A0 = __builtin_fmaf(-q1, y0, F.A0) + o0;
    F.A1 = __builtin_fmaf(-q1, y1, F.A1) + o1;
    F.A2 = __builtin_fmaf(-q1, y2, F.A2) + o2;
    F.S0 = __builtin_fmaf(c2, u0, F.S0);
    F.S1 = __builtin_fmaf(c2, u1, F.S1);
    F.S2 = __builtin_fmaf(c2, u2, F.S2);
    const float w0 = c1 * u0, w1 = c1 * u1, w2 = c1 * u2;
    F.Pxx = __builtin_fmaf(w0, u0, F.Pxx);
    F.Pyy = __builtin_fmaf(w1, u1, F.Pyy);
    F.Pzz = __builtin_fmaf(w2, u2, F.Pzz);
    F.Pxy = __builtin_fmaf(w0, u1, F.Pxy);
    F.Pxz = __builtin_fmaf(w0, u2, F.Pxz);
    F.Pyz = __builtin_fmaf(w1, u2, F.Pyz);
    const float v0 = c3 * u0, v1 = c3 * u1, v2 = c3 * u2;
    const float v00 = v0 * u0, v01 = v0 * u1, v02 = v0 * u2, v11 = v1 * u1, v12 = v1 * u2,
                v22 = v2 * u2;
    F.Txxx = __builtin_fmaf(v00, u0, F.Txxx);
    F.Txxy = __builtin_fmaf(v00, u1, F.Txxy);
    F.Txxz = __builtin_fmaf(v00, u2, F.Txxz);
    F.Txyy = __builtin_fmaf(v01, u1, F.Txyy);
    F.Txyz = __builtin_fmaf(v01, u2, F.Txyz);
    F.Txzz = __builtin_fmaf(v02, u2, F.Txzz);
    F.Tyyy = __builtin_fmaf(v11, u1, F.Tyyy);
    F.Tyyz = __builtin_fmaf(v11, u2, F.Tyyz);
    F.Tyzz = __builtin_fmaf(v12, u2, F.Tyzz);
    F.Tzzz = __builtin_fmaf(v22, u2, F.Tzzz);
}

__device__ __forceinline__ float wave_sumf(float v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Wave totals of the moments, then the expansion at offset d from the centre:
//   a_k = a0_k + (P d)_k - Q d_k + d_k (S.d) + (1/2) S_k |d|^2 - (1/2) (T d d)_k
__device__ __forceinline__ void far_eval(FarMoments F, float d0, float d1, float d2, float &o0,
                                         float &o1, float &o2)
{
    F.Q = wave_sumf(F.Q);
    F.A0 = wave_sumf(F.A0), F.A1 = wave_sumf(F.A1), F.A2 = wave_sumf(F.A2);
    F.S0 = wave_sumf(F.S0), F.S1 = wave_sumf(F.S1), F.S2 = wave_sumf(F.S2);
    F.Pxx = wave_sumf(F.Pxx), F.Pyy = wave_sumf(F.Pyy), F.Pzz = wave_sumf(F.Pzz);
    F.Pxy = wave_sumf(F.Pxy), F.Pxz = wave_sumf(F.Pxz), F.Pyz = wave_sumf(F.Pyz);
    F.Txxx = wave_sumf(F.Txxx), F.Tyyy = wave_sumf(F.Tyyy), F.Tzzz = wave_sumf(F.Tzzz);
    F.Txxy = wave_sumf(F.Txxy), F.Txxz = wave_sumf(F.Txxz), F.Txyy = wave_sumf(F.Txyy);
    F.Tyyz = wave_sumf(F.Tyyz), F.Txzz = wave_sumf(F.Txzz), F.Tyzz = wave_sumf(F.Tyzz);
    F.Txyz = wave_sumf(F.Txyz);
    const float Sd = __builtin_fmaf(F.S2, d2, __builtin_fmaf(F.S1, d1, F.S0 * d0));
    const float dd = __builtin_fmaf(d2, d2, __builtin_fmaf(d1, d1, d0 * d0));
    const float x2 = d0 * d0, y2 = d1 * d1, z2 = d2 * d2;
    const float xy = 2.0f * d0 * d1, xz = 2.0f * d0 * d2, yz = 2.0f * d1 * d2;
    const float T0 = F.Txxx * x2 + F.Txyy * y2 + F.Txzz * z2 + F.Txxy * xy + F.Txxz * xz + F.Txyz * yz;
    const float T1 = F.Txxy * x2 + F.Tyyy * y2 + F.Tyzz * z2 + F.Txyy * xy + F.Txyz * xz + F.Tyyz * yz;
    const float T2 = F.Txxz * x2 + F.Tyyz * y2 + F.Tzzz * z2 + F.Txyz * xy + F.Txzz * xz + F.Tyzz * yz;
    const float P0 = F.Pxx * d0 + F.Pxy * d1 + F.Pxz * d2;
    const float P1 = F.Pxy * d0 + F.Pyy * d1 + F.Pyz * d2;
    const float P2 = F.Pxz * d0 + F.Pyz * d1 + F.Pzz * d2;
    o0 = F.A0 + (P0 - F.Q * d0) + d0 * Sd + 0.5f * (F.S0 * dd - T0);
    o1 = F.A1 + (P1 - F.Q * d1) + d1 * Sd + 0.5f * (F.S1 * dd - T1);
    o2 = F.A2 + (P2 - F.Q * d2) + d2 * Sd + 0.5f * (F.S2 * dd - T2);
}

template <bool QUAD, bool LE>
__global__ __launch_bounds__(GW_BLOCK) void k_walk_group(const NodeRec *__restrict__ nodes,
                                                         const double4 *__restrict__ P,
                                                         const int *__restrict__ sidx, int n,
                                                         const int2 *__restrict__ groups,
                                                         int num_groups, WalkArgs args,
                                                         double *__restrict__ a,
                                                         unsigned long long *__restrict__ n_int,
                                                         const float *__restrict__ quad)
{
    __shared__ int stk_all[GW_WAVES][GW_STACK];
    __shared__ float4 nl_all[GW_WAVES][GW_LIST];  // accepted nodes: com - centre, mass
    __shared__ int ni_all[GW_WAVES][GW_LIST];     // their node indices
    __shared__ float4 pl_all[GW_WAVES][GW_LIST];  // particles: x - centre, mass
    __shared__ int pi_all[GW_WAVES][GW_LIST];     // their sorted positions
    __shared__ int rinc_all[GW_WAVES][64];        // opened multi-particle leaves:
    __shared__ int rbeg_all[GW_WAVES][64];        // inclusive count prefix, first - offset
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int gi = blockIdx.x * GW_WAVES + wave;
    if (gi >= num_groups) return;
    const int2 grp = groups[gi];  // (first sorted position, count <= 64)
    const int g0 = grp.x;
    const int gend = g0 + grp.y;
    const int s = g0 + lane;
    const bool valid = s < gend;
    int *stk = stk_all[wave];
    float4 *nl = nl_all[wave];
    int *ni = ni_all[wave];
    float4 *pl = pl_all[wave];
    int *pi = pi_all[wave];
    int *rinc = rinc_all[wave];
    int *rbeg = rbeg_all[wave];

    const double4 ps = P[valid ? s : g0];
    const double lo0 = uniform(wave_min(ps.x)), lo1 = uniform(wave_min(ps.y)),
                 lo2 = uniform(wave_min(ps.z));
    const double hi0 = uniform(wave_max(ps.x)), hi1 = uniform(wave_max(ps.y)),
                 hi2 = uniform(wave_max(ps.z));
    const double gc0 = 0.5 * (lo0 + hi0), gc1 = 0.5 * (lo1 + hi1), gc2 = 0.5 * (lo2 + hi2);
    const float xr = static_cast<float>(ps.x - gc0), yr = static_cast<float>(ps.y - gc1),
                zr = static_cast<float>(ps.z - gc2);
    const float eps2 = static_cast<float>(args.eps2);
    // Particle pairs closer than 2^-7 of the group's width: their FP32
    // coordinates relative to the group centre carry an error of ~2^-25 of
    // that width each, so such a pair's separation is formed again in FP64
    // from the particle records (to about 2^-17 of the separation otherwise).
    const float gw = static_cast<float>(fmax(fmax(hi0 - lo0, hi1 - lo1), hi2 - lo2));
    const float close2 = __builtin_fmaf(gw * 0x1p-7f, gw * 0x1p-7f, eps2);
    // local expansions: an accepted node is far when rg < phi |com - c|, rg
    // the half-diagonal of the group's box
    const double hb0 = 0.5 * (hi0 - lo0), hb1 = 0.5 * (hi1 - lo1), hb2 = 0.5 * (hi2 - lo2);
    const double rg2 = hb0 * hb0 + hb1 * hb1 + hb2 * hb2;
    FarMoments F = {};
    unsigned nfar = 0;
    double A0 = 0.0, A1 = 0.0, A2 = 0.0;
    unsigned cnt = 0;
    bool overflow = false;
    int top = 1, nn = 0, np = 0;

    // Sum of list entries [0, m) for this lane's particle.
    auto sum_nodes = [&](int m) {
        float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
        for (int e = 0; e < m; e++) {
            const float4 q = nl[e];
            const float dx = q.x - xr, dy = q.y - yr, dz = q.z - zr;
            float r2 = __builtin_fmaf(dx, dx, eps2);
            r2 = __builtin_fmaf(dy, dy, r2);
            r2 = __builtin_fmaf(dz, dz, r2);
            const float ri = __builtin_amdgcn_rsqf(r2);
            const float w = (q.w * ri) * (ri * ri);
            f0 = __builtin_fmaf(dx, w, f0);
            f1 = __builtin_fmaf(dy, w, f1);
            f2 = __builtin_fmaf(dz, w, f2);
            if (QUAD) {
                const int node = __builtin_amdgcn_readfirstlane(ni[e]);
                float o0, o1, o2;
                quad_termf(quad + 8 * static_cast<size_t>(node), -dx, -dy, -dz, ri, o0, o1, o2);
                f0 += o0;
                f1 += o1;
                f2 += o2;
            }
        }
        A0 += static_cast<double>(f0);
        A1 += static_cast<double>(f1);
        A2 += static_cast<double>(f2);
        cnt += m;
    };
    auto sum_parts = [&](int m) {
        float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
        float mn = __builtin_inff();
        for (int e = 0; e < m; e++) {
            const float4 q = pl[e];
            const float dx = q.x - xr, dy = q.y - yr, dz = q.z - zr;
            float r2 = __builtin_fmaf(dx, dx, eps2);
            r2 = __builtin_fmaf(dy, dy, r2);
            r2 = __builtin_fmaf(dz, dz, r2);
            const float ri = __builtin_amdgcn_rsqf(r2);
            const bool me = pi[e] == s;
            const float w = me ? 0.0f : (q.w * ri) * (ri * ri);
            mn = fminf(mn, me ? __builtin_inff() : r2);
            f0 = __builtin_fmaf(dx, w, f0);
            f1 = __builtin_fmaf(dy, w, f1);
            f2 = __builtin_fmaf(dz, w, f2);
        }
        if (__ballot(valid && mn < close2)) {
            // a close pair in this batch: sum it again with those separations
            // taken from the FP64 records
            f0 = f1 = f2 = 0.0f;
            for (int e = 0; e < m; e++) {
                const float4 q = pl[e];
                const int pe = pi[e];
                const bool me = pe == s;
                float dx = q.x - xr, dy = q.y - yr, dz = q.z - zr;
                float r2 = __builtin_fmaf(dx, dx, eps2);
                r2 = __builtin_fmaf(dy, dy, r2);
                r2 = __builtin_fmaf(dz, dz, r2);
                if (valid && !me && r2 < close2) {
                    const double4 pp = P[pe];
                    dx = static_cast<float>(pp.x - ps.x);
                    dy = static_cast<float>(pp.y - ps.y);
                    dz = static_cast<float>(pp.z - ps.z);
                    r2 = __builtin_fmaf(dx, dx, eps2);
                    r2 = __builtin_fmaf(dy, dy, r2);
                    r2 = __builtin_fmaf(dz, dz, r2);
                }
                const float ri = __builtin_amdgcn_rsqf(r2);
                const float w = me ? 0.0f : (q.w * ri) * (ri * ri);
                f0 = __builtin_fmaf(dx, w, f0);
                f1 = __builtin_fmaf(dy, w, f1);
                f2 = __builtin_fmaf(dz, w, f2);
            }
        }
        A0 += static_cast<double>(f0);
        A1 += static_cast<double>(f1);
        A2 += static_cast<double>(f2);
        cnt += m;
    };

    // sum the first 64 particle-list entries, move the rest (< 64) to the front
    auto flush_parts = [&]() {
        sum_parts(64);
        __builtin_amdgcn_wave_barrier();
        if (lane < np - 64) {
            pl[lane] = pl[64 + lane];
            pi[lane] = pi[64 + lane];
        }
        np -= 64;
        __builtin_amdgcn_wave_barrier();
    };

    // from the root itself: it holds the group, so it is opened (a root that
    // is a leaf -- n within the leaf capacity -- gives its particles directly)
    // Each pass pops the next pass's entries and requests their records as
    // soon as its own children are on the stack, so that the loads overlap
    // the list sums.  The first pass is the root alone.
    top = 0;
    int k = 1;
    int node = 0;
    NodeRec nd = nodes[0];
    while (k > 0) {
        const bool act = lane < k;
        const bool leaf = nd.nchild == 0;
        const bool single_leaf = act && leaf && nd.count == 1;
        const bool incl = nd.first < gend && nd.first + nd.count > g0;
        const double e0 = fmax(fmax(lo0 - nd.cx, nd.cx - hi0), 0.0);
        const double e1 = fmax(fmax(lo1 - nd.cy, nd.cy - hi1), 0.0);
        const double e2 = fmax(fmax(lo2 - nd.cz, nd.cz - hi2), 0.0);
        const double d2 = e0 * e0 + e1 * e1 + e2 * e2;
        bool accept = !incl && nd.l2 < args.theta2 * d2;
        bool far = false;
        if (LE) {
            // far nodes (single-particle leaves too), accepted at the far
            // opening angle: their moments about the group centre, one node
            // per lane
            const double y0 = gc0 - nd.cx, y1 = gc1 - nd.cy, y2 = gc2 - nd.cz;
            const bool farz = rg2 < args.phi2 * (y0 * y0 + y1 * y1 + y2 * y2);
            accept = accept || (!incl && farz && nd.l2 < args.thetaf2 * d2);
            far = act && accept && farz;
            if (far)
                far_add(F, static_cast<float>(y0), static_cast<float>(y1), static_cast<float>(y2),
                        nd.fmass, eps2, quad + 8 * static_cast<size_t>(node));
            const unsigned nf = __popcll(__ballot(far));
            nfar += nf;
            cnt += nf;
        }
        // single-particle leaves not expanded: their particle into the particle list
        const bool single = single_leaf && !far;
        const bool take_node = act && !single_leaf && accept && !far;
        const bool multi = act && leaf && !single_leaf && !accept;
        const bool open = act && !leaf && !accept;
        // accepted nodes -> node list
        {
            const unsigned long long B = __ballot(take_node);
            if (take_node) {
                const int at = nn + __popcll(B & ((1ull << lane) - 1));
                nl[at] = make_float4(static_cast<float>(nd.cx - gc0), static_cast<float>(nd.cy - gc1),
                                     static_cast<float>(nd.cz - gc2), nd.fmass);
                if (QUAD) ni[at] = node;
            }
            nn += __popcll(B);
        }
        // single-particle leaves -> particle list (the particle itself)
        {
            const unsigned long long B = __ballot(single);
            if (single) {
                const int at = np + __popcll(B & ((1ull << lane) - 1));
                const double4 p = P[nd.first];
                pl[at] = make_float4(static_cast<float>(p.x - gc0), static_cast<float>(p.y - gc1),
                                     static_cast<float>(p.z - gc2), static_cast<float>(p.w));
                pi[at] = nd.first;
            }
            np += __popcll(B);
        }
        // opened internal nodes -> children on the stack, in lane order: the
        // exclusive prefix of the child counts (0..8) from four bit-plane
        // ballots, without a cross-lane shuffle chain
        {
            const int c = open ? nd.nchild : 0;
            const unsigned long long below = (1ull << lane) - 1;
            int before = 0, total = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const unsigned long long B = __ballot((c >> b) & 1);
                before += __popcll(B & below) << b;
                total += __popcll(B) << b;
            }
            if (top + total <= GW_STACK) {
                for (int q = 0; q < c; q++) stk[top + before + q] = nd.child0 + q;
                top += total;
            } else {
                overflow = true;
            }
        }
        __builtin_amdgcn_wave_barrier();
        // the next pass: at most 64 entries, fewer when its children could
        // bring the stack past the batch limit
        int kn = min(top, 64);
        if (top + 7 * kn > GW_BATCH_LIMIT) kn = max(1, (GW_BATCH_LIMIT - top) / 7);
        kn = min(kn, top);
        const int node_next = lane < kn ? stk[top - kn + lane] : 0;
        top -= kn;
        __builtin_amdgcn_wave_barrier();
        const NodeRec nd_next = nodes[node_next];
        if (np >= 64) flush_parts();
        // opened leaves of several particles: their particles into the
        // particle list, 64 slots at a time, every lane loading one particle
        // (slot t belongs to the first multi leaf whose inclusive count
        // prefix exceeds t)
        if (__ballot(multi)) {
            const int c = multi ? nd.count : 0;
            int inc = c;
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(inc, o);
                if (lane >= o) inc += t;
            }
            const int total = __shfl(inc, 63);
            if (total > 0) {
                rinc[lane] = inc;
                rbeg[lane] = nd.first - (inc - c);  // particle of slot t: rbeg[L] + t
                __builtin_amdgcn_wave_barrier();
                for (int j0 = 0; j0 < total; j0 += 64) {
                    const int m = min(64, total - j0);
                    if (lane < m) {
                        const int t = j0 + lane;
                        int L = 0;
                        for (int step = 32; step > 0; step >>= 1)
                            if (rinc[L + step - 1] <= t) L += step;
                        const int q = rbeg[L] + t;
                        const double4 p = P[q];
                        pl[np + lane] = make_float4(static_cast<float>(p.x - gc0), static_cast<float>(p.y - gc1),
                                                    static_cast<float>(p.z - gc2), static_cast<float>(p.w));
                        pi[np + lane] = q;
                    }
                    np += m;
                    __builtin_amdgcn_wave_barrier();
                    if (np >= 64) flush_parts();
                }
            }
        }
        // full lists: sum 64 entries, move the rest (< 64) to the front
        if (nn >= 64) {
            sum_nodes(64);
            __builtin_amdgcn_wave_barrier();
            if (lane < nn - 64) {
                nl[lane] = nl[64 + lane];
                if (QUAD) ni[lane] = ni[64 + lane];
            }
            nn -= 64;
            __builtin_amdgcn_wave_barrier();
        }
        nd = nd_next;
        node = node_next;
        k = kn;
    }
    if (nn > 0) sum_nodes(nn);
    if (np > 0) sum_parts(np);
    if (LE && nfar > 0) {
        float o0, o1, o2;
        far_eval(F, xr, yr, zr, o0, o1, o2);
        A0 += static_cast<double>(o0);
        A1 += static_cast<double>(o1);
        A2 += static_cast<double>(o2);
    }
    if (valid) {
        const int i = sidx[s];
        a[3 * i + 0] = args.G * A0;
        a[3 * i + 1] = args.G * A1;
        a[3 * i + 2] = args.G * A2;
    }
    // (a particle's own entry is in its group's particle list exactly once:
    // every node holding it is opened, down to its leaf)
    unsigned long long tot = valid ? cnt - 1u : 0u;
    for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
    if (lane == 0 && n_int) {
        atomicAdd(n_int, tot + (overflow ? (1ull << 62) : 0ull));
        if (LE) atomicAdd(n_int + 1, static_cast<unsigned long long>(nfar));
    }
}

// ------------------------------------------------------------------ host side
static DeviceTree g_tree;

static ErrorStatus ensure_nodes(Context *c, DeviceTree *t, int need)
{
    if (need <= t->node_cap) return make_success_error_status();
    int cap = t->node_cap > 0 ? t->node_cap : 1024;
    while (cap < need) cap *= 2;
    GS_HIP(hipStreamSynchronize(c->stream));
    NodeRec *nn = nullptr;
    GS_HIP(hipMalloc(reinterpret_cast<void **>(&nn), sizeof(NodeRec) * cap));
    if (t->nodes) {
        GS_HIP(hipMemcpy(nn, t->nodes, sizeof(NodeRec) * t->node_cap, hipMemcpyDeviceToDevice));
        GS_HIP(hipFree(t->nodes));
    }
    t->nodes = nn;
    double4 *nm = nullptr;
    GS_HIP(hipMalloc(reinterpret_cast<void **>(&nm), sizeof(double4) * cap));
    if (t->mom) GS_HIP(hipFree(t->mom));
    t->mom = nm;
    double *nq = nullptr;
    GS_HIP(hipMalloc(reinterpret_cast<void **>(&nq), sizeof(double) * 8 * cap));
    if (t->quad) GS_HIP(hipFree(t->quad));
    t->quad = nq;
    float *nqf = nullptr;
    GS_HIP(hipMalloc(reinterpret_cast<void **>(&nqf), sizeof(float) * 8 * cap));
    if (t->quadf) GS_HIP(hipFree(t->quadf));
    t->quadf = nqf;
    unsigned *tc = nullptr;
    GS_HIP(hipMalloc(reinterpret_cast<void **>(&tc), sizeof(unsigned) * (cap + 1)));
    if (t->tmp) GS_HIP(hipFree(t->tmp));
    t->tmp = tc;
    void *st = nullptr;
    GS_HIP(hipMalloc(&st, scan_temp_bytes(cap) + 64));
    if (t->scan_tmp) GS_HIP(hipFree(t->scan_tmp));
    t->scan_tmp = static_cast<unsigned *>(st);
    t->node_cap = cap;
    return make_success_error_status();
}

size_t radix_hist_elements(int n)
{
    // the tile counts, then the digit totals
    return static_cast<size_t>((n + RS_TILE - 1) / RS_TILE) * RS_BINS + RS_BINS;
}

ErrorStatus radix_sort_pairs(Context *c, long long *k0, int *i0, long long *k1, int *i1, int n,
                             int passes, unsigned *hist, long long **ko, int **io)
{
    const int ntiles = (n + RS_TILE - 1) / RS_TILE;
    unsigned *dtot = hist + static_cast<size_t>(ntiles) * RS_BINS;
    for (int pass = 0; pass < passes; pass++) {
        const int shift = pass * RS_BITS;
        hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(RS_BLOCK), 0, c->stream, k0, n, shift,
                           hist, ntiles);
        hipLaunchKernelGGL(k_rs_rowscan, dim3(RS_BINS), dim3(RS_BLOCK), 0, c->stream, hist, ntiles,
                           dtot);
        hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(RS_BLOCK), 0, c->stream, k0, i0, n,
                           shift, hist, dtot, ntiles, k1, i1);
        GS_HIP(hipGetLastError());
        long long *tk = k0;
        k0 = k1;
        k1 = tk;
        int *ti = i0;
        i0 = i1;
        i1 = ti;
    }
    *ko = k0;
    *io = i0;
    return make_success_error_status();
}

ErrorStatus tree_sort(Context *c, const double *x, int n, const double *box_center,
                      double box_width, DeviceTree *t, const double *m, double4 *packed)
{
    if (!m) packed = nullptr;
    if (n <= 0) return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR, "Number of particles must be positive");
    void *p;
    GS_TRY(slot(c, SLOT_BBOX, 64 * sizeof(double) + 6 * 1024 * sizeof(double), &p));
    double *box = static_cast<double *>(p);
    double *bpart = box + 64;
    GS_TRY(slot(c, SLOT_KEYS, sizeof(long long) * n, &p));
    long long *k0 = static_cast<long long *>(p);
    GS_TRY(slot(c, SLOT_KEYS_ALT, sizeof(long long) * n, &p));
    long long *k1 = static_cast<long long *>(p);
    GS_TRY(slot(c, SLOT_IDX, sizeof(int) * n, &p));
    int *i0 = static_cast<int *>(p);
    GS_TRY(slot(c, SLOT_IDX_ALT, sizeof(int) * n, &p));
    int *i1 = static_cast<int *>(p);
    GS_TRY(slot(c, SLOT_HIST, sizeof(unsigned) * radix_hist_elements(n), &p));
    unsigned *hist = static_cast<unsigned *>(p);

    const int rec = timing_begin(c, "tree_sort");
    if (box_center && box_width > 0.0) {
        double hb[4] = {box_center[0], box_center[1], box_center[2], box_width};
        GS_HIP(hipMemcpyAsync(box, hb, sizeof(hb), hipMemcpyHostToDevice, c->stream));
    } else {
        int nb = (n + BB_BLOCK - 1) / BB_BLOCK;
        if (nb > 1024) nb = 1024;
        hipLaunchKernelGGL(k_bbox_partial, dim3(nb), dim3(BB_BLOCK), 0, c->stream, x, n, bpart);
        hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(BB_BLOCK), 0, c->stream, bpart, nb, box);
    }
    hipLaunchKernelGGL(k_morton, dim3((n + 255) / 256), dim3(256), 0, c->stream, x, n, box, k0,
                       i0, m, packed);
    GS_HIP(hipGetLastError());
    long long *ks;
    int *is;
    GS_TRY(radix_sort_pairs(c, k0, i0, k1, i1, n, RS_PASSES, hist, &ks, &is));
    k0 = ks;
    i0 = is;
    timing_end(c, rec);
    // 8 passes: the sorted data is back in the primary slots
    t->n = n;
    t->skeys = k0;
    t->sidx = i0;
    t->box = box;
    double hb[4];
    GS_HIP(hipMemcpyAsync(hb, box, sizeof(hb), hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    for (int d = 0; d < 3; d++) t->center[d] = hb[d];
    t->width = hb[3];
    return make_success_error_status();
}

// Moments bottom-up, one launch per level (deepest first).  Third moments
// (with_oct, which implies with_quad) in arrays allocated on first use.
static ErrorStatus tree_moments(Context *c, DeviceTree *t, bool with_quad, bool with_oct)
{
    if (with_oct && t->oct_cap < t->node_cap) {
        GS_HIP(hipStreamSynchronize(c->stream));
        if (t->oct) GS_HIP(hipFree(t->oct));
        if (t->octf) GS_HIP(hipFree(t->octf));
        t->oct = nullptr;
        t->octf = nullptr;
        t->oct_cap = 0;
        GS_HIP(hipMalloc(reinterpret_cast<void **>(&t->oct), sizeof(double) * 12 * t->node_cap));
        GS_HIP(hipMalloc(reinterpret_cast<void **>(&t->octf), sizeof(float) * 12 * t->node_cap));
        t->oct_cap = t->node_cap;
    }
    void *gp;
    GS_TRY(slot(c, SLOT_GROUP_FLAGS, sizeof(unsigned) * static_cast<size_t>(t->num_nodes) + 256, &gp));
    t->gflag = static_cast<unsigned *>(gp);
    for (int d = t->num_levels - 1; d >= 0; d--) {
        const int lb = t->level_begin[d], le = t->level_end[d];
        if (le <= lb) continue;
        const dim3 grid((le - lb + 255) / 256);
        if (with_oct)
            hipLaunchKernelGGL((k_multipole<true, true>), grid, dim3(256), 0, c->stream, t->nodes, lb,
                               le, t->P, t->mom, t->quad, t->quadf, t->oct, t->octf, t->gflag);
        else if (with_quad)
            hipLaunchKernelGGL((k_multipole<true, false>), grid, dim3(256), 0, c->stream, t->nodes,
                               lb, le, t->P, t->mom, t->quad, t->quadf, t->oct, t->octf, t->gflag);
        else
            hipLaunchKernelGGL((k_multipole<false, false>), grid, dim3(256), 0, c->stream, t->nodes,
                               lb, le, t->P, t->mom, t->quad, t->quadf, t->oct, t->octf, t->gflag);
    }
    GS_HIP(hipGetLastError());
    return make_success_error_status();
}

// All levels at once (leaf_cap <= FB_MAX_LEAF): digits and leaf depths per
// particle, node counts per (depth, tile), one scan and the emission; one
// host round trip (the level sizes), then the moments level by level.
static ErrorStatus tree_build_flat(Context *c, DeviceTree *t, int n, int leaf_cap, bool with_quad,
                                   bool with_oct)
{
    const int ntiles = (n + FB_TILE - 1) / FB_TILE;
    const long long hn = static_cast<long long>(FB_DEPTHS) * ntiles;
    if (hn > scan_max_elements())
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "Too many particles for the octree: %d", n);
    auto up = [](size_t b) { return (b + 255) & ~static_cast<size_t>(255); };
    const size_t bL = up(static_cast<size_t>(n) + 1 + FB_PAD), bD = up(n),
                 bC = up(sizeof(unsigned) * static_cast<size_t>(hn)), bV = up(sizeof(unsigned) * 32);
    void *p;
    GS_TRY(slot(c, SLOT_LEVELS, bL + bD + bC + bV, &p));
    signed char *Ld = static_cast<signed char *>(p);
    signed char *Dd = Ld + bL;
    unsigned *cnt = reinterpret_cast<unsigned *>(Dd + bD);
    unsigned *lev = reinterpret_cast<unsigned *>(reinterpret_cast<char *>(cnt) + bC);
    GS_TRY(slot(c, SLOT_HIST_SCAN, scan_temp_bytes(hn) + 64, &p));
    unsigned *stmp = static_cast<unsigned *>(p);
    GS_TRY(slot(c, SLOT_MISC, 256, &p));
    unsigned *total_dev = static_cast<unsigned *>(p);
    void *hp;
    GS_TRY(pinned(c, 256, &hp));
    unsigned *lev_host = static_cast<unsigned *>(hp);

    hipLaunchKernelGGL(k_fb_digits, dim3((n + 1 + FB_PAD + 255) / 256), dim3(256), 0, c->stream,
                       t->skeys, n, Ld);
    hipLaunchKernelGGL(k_fb_leafdepth, dim3((n + 255) / 256), dim3(256), 0, c->stream, Ld, n,
                       leaf_cap, Dd);
    hipLaunchKernelGGL(k_fb_count, dim3(ntiles), dim3(FB_BLOCK), 0, c->stream, Ld, Dd, n, ntiles,
                       cnt);
    GS_HIP(hipGetLastError());
    GS_HIP(scan_exclusive(c->stream, cnt, cnt, hn, stmp, total_dev));
    hipLaunchKernelGGL(k_fb_levels, dim3(1), dim3(64), 0, c->stream, cnt, ntiles, total_dev, lev);
    GS_HIP(hipMemcpyAsync(lev_host, lev, sizeof(unsigned) * (FB_DEPTHS + 1), hipMemcpyDeviceToHost,
                          c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    const int nn = static_cast<int>(lev_host[FB_DEPTHS]);
    GS_TRY(ensure_nodes(c, t, nn));
    // per-node flags of the cells left open, in the child-count scratch (which
    // only the level-by-level build uses; cap + 1 words >= nn bytes)
    unsigned char *fix = reinterpret_cast<unsigned char *>(t->tmp);
    hipLaunchKernelGGL(k_fb_emit, dim3(ntiles), dim3(FB_BLOCK), 0, c->stream, t->skeys, Ld, Dd, n,
                       ntiles, cnt, t->width * t->width, t->nodes, fix);
    hipLaunchKernelGGL(k_fb_fix, dim3((nn + 255) / 256), dim3(256), 0, c->stream, t->nodes, nn,
                       t->skeys, n, fix, lev);
    GS_HIP(hipGetLastError());
    int levels = 0;
    for (int d = 0; d < FB_DEPTHS; d++) {
        const int b = static_cast<int>(lev_host[d]);
        const int e = d + 1 < FB_DEPTHS ? static_cast<int>(lev_host[d + 1]) : nn;
        t->level_begin[d] = b;
        t->level_end[d] = e;
        if (e > b) levels = d + 1;
    }
    t->num_levels = levels;
    t->num_nodes = nn;
    return tree_moments(c, t, with_quad, with_oct);
}

ErrorStatus tree_build(Context *c, const double *x, const double *m, int n, int leaf_cap,
                       const double *box_center, double box_width, bool with_quad,
                       DeviceTree **out, bool with_oct)
{
    with_quad = with_quad || with_oct;
    DeviceTree *t = &g_tree;
    t->has_quad = false;
    t->has_oct = false;
    // an unfinalised parameter (-1) means the reference's default of one
    // particle per leaf (src/acceleration.c:73-129)
    if (leaf_cap < 1) leaf_cap = 1;
    void *p;
    GS_TRY(slot(c, SLOT_PACKED_P, sizeof(double4) * n, &p));
    double4 *packed = static_cast<double4 *>(p);
    GS_TRY(tree_sort(c, x, n, box_center, box_width, t, m, packed));
    t->leaf_cap = leaf_cap;
    GS_TRY(slot(c, SLOT_SORTED_P, sizeof(double4) * n, &p));
    t->P = static_cast<double4 *>(p);
    hipLaunchKernelGGL(k_gather_particles, dim3((n + 255) / 256), dim3(256), 0, c->stream, packed,
                       t->sidx, n, t->P);
    GS_HIP(hipGetLastError());

    const int rec = timing_begin(c, "tree_build");
    if (leaf_cap <= FB_MAX_LEAF) {
        GS_TRY(tree_build_flat(c, t, n, leaf_cap, with_quad, with_oct));
        t->has_quad = with_quad;
        t->has_oct = with_oct;
        timing_end(c, rec);
        *out = t;
        return make_success_error_status();
    }
    GS_TRY(ensure_nodes(c, t, 2 * n + 64));
    NodeRec root;
    memset(&root, 0, sizeof(root));
    root.depth = 0;
    root.prefix = 0;
    root.first = 0;
    root.count = n;
    root.child0 = -1;
    root.l2 = t->width * t->width;
    GS_HIP(hipMemcpyAsync(t->nodes, &root, sizeof(root), hipMemcpyHostToDevice, c->stream));
    t->level_begin[0] = 0;
    t->level_end[0] = 1;
    int levels = 1;
    unsigned *total_dev = nullptr;
    GS_TRY(slot(c, SLOT_MISC, 256, &p));
    total_dev = static_cast<unsigned *>(p);
    void *hp;
    GS_TRY(pinned(c, 256, &hp));
    unsigned *total_host = static_cast<unsigned *>(hp);
    for (int d = 1; d <= MAXL; d++) {
        const int lb = t->level_begin[d - 1], le = t->level_end[d - 1];
        const int np = le - lb;
        if (np <= 0) break;
        const int grid = (np + 255) / 256;
        hipLaunchKernelGGL(k_count_children, dim3(grid), dim3(256), 0, c->stream, t->nodes, lb,
                           le, leaf_cap, t->skeys, t->tmp);
        GS_HIP(scan_exclusive(c->stream, t->tmp, t->tmp, np, t->scan_tmp, total_dev));
        GS_HIP(hipMemcpyAsync(total_host, total_dev, sizeof(unsigned), hipMemcpyDeviceToHost,
                              c->stream));
        GS_HIP(hipStreamSynchronize(c->stream));
        const int nc = static_cast<int>(*total_host);
        if (le + nc > t->node_cap) {
            // grow (keeps the node array) and recount into the new scratch
            GS_TRY(ensure_nodes(c, t, le + nc));
            hipLaunchKernelGGL(k_count_children, dim3(grid), dim3(256), 0, c->stream, t->nodes,
                               lb, le, leaf_cap, t->skeys, t->tmp);
            GS_HIP(scan_exclusive(c->stream, t->tmp, t->tmp, np, t->scan_tmp, total_dev));
        }
        hipLaunchKernelGGL(k_emit_children, dim3(grid), dim3(256), 0, c->stream, t->nodes, lb, le,
                           leaf_cap, t->width * t->width, t->skeys, t->tmp, t->node_cap);
        GS_HIP(hipGetLastError());
        t->level_begin[d] = le;
        t->level_end[d] = le + nc;
        levels = d + 1;
        if (nc == 0) break;
    }
    // the deepest emitted level has no children (depth 21 is always a leaf)
    {
        const int d = levels - 1;
        const int lb = t->level_begin[d], le = t->level_end[d];
        if (le > lb && d == MAXL) {
            hipLaunchKernelGGL(k_emit_children, dim3((le - lb + 255) / 256), dim3(256), 0,
                               c->stream, t->nodes, lb, le, leaf_cap, t->width * t->width,
                               t->skeys, t->tmp, t->node_cap);
        }
    }
    t->num_levels = levels;
    t->num_nodes = t->level_end[levels - 1];
    GS_TRY(tree_moments(c, t, with_quad, with_oct));
    GS_HIP(hipGetLastError());
    t->has_quad = with_quad;
    t->has_oct = with_oct;
    timing_end(c, rec);
    *out = t;
    return make_success_error_status();
}

ErrorStatus tree_walk(Context *c, DeviceTree *t, const double *x, double G, double eps,
                      double theta, double *a, long long *interactions, int order, float *feat)
{
    (void)x;
    if ((order == 2 || feat) && !t->has_quad)
        return WRAP_RAISE_ERROR(GRAV_FAILURE, "Quadrupole or feature walk of a tree built without "
                                              "second moments");
    if (feat && !t->has_oct)
        return WRAP_RAISE_ERROR(GRAV_FAILURE, "Feature walk of a tree built without third moments");
    WalkArgs args;
    args.theta2 = theta * theta;
    args.eps2 = eps * eps;
    args.G = G;
    args.phi2 = 0.0;
    args.thetaf2 = args.theta2;
    void *p;
    GS_TRY(slot(c, SLOT_MISC, 256, &p));
    unsigned long long *cnt = reinterpret_cast<unsigned long long *>(static_cast<char *>(p) + 64);
    GS_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), c->stream));
    const int waves = (t->n + 63) / 64;
    const int grid = (waves + WK_WAVES - 1) / WK_WAVES;
    const int rec = timing_begin(c, "bh_walk");
    if (feat)
        hipLaunchKernelGGL((k_walk<false, true>), dim3(grid), dim3(WK_BLOCK), 0, c->stream,
                           t->nodes, t->P, t->skeys, t->sidx, t->n, args, a, cnt, t->quadf, t->octf,
                           feat);
    else if (order >= 2)
        hipLaunchKernelGGL((k_walk<true, false>), dim3(grid), dim3(WK_BLOCK), 0, c->stream,
                           t->nodes, t->P, t->skeys, t->sidx, t->n, args, a, cnt, t->quadf, t->octf,
                           feat);
    else
        hipLaunchKernelGGL((k_walk<false, false>), dim3(grid), dim3(WK_BLOCK), 0, c->stream,
                           t->nodes, t->P, t->skeys, t->sidx, t->n, args, a, cnt, t->quadf, t->octf,
                           feat);
    timing_end(c, rec);
    GS_HIP(hipGetLastError());
    if (interactions) {
        unsigned long long h = 0;
        GS_HIP(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        GS_HIP(hipStreamSynchronize(c->stream));
        if (h >> 62)
            return WRAP_RAISE_ERROR(GRAV_FAILURE, "Barnes-Hut walk stack overflow");
        *interactions = static_cast<long long>(h);
    }
    return make_success_error_status();
}

// Group walk (methods 20 and 21): order 1 monopoles, 2 with quadrupoles.
// Widest box of a walk group before it is split into single particles: four
// times the width of a 64-particle group at the mean density of the tree's box.
static double group_width_cap(double width, int n)
{
    return 4.0 * width * std::cbrt(64.0 / static_cast<double>(n));
}

ErrorStatus tree_walk_group(Context *c, DeviceTree *t, double G, double eps, double theta,
                            double *a, long long *interactions, int order, double phi,
                            double theta_far, long long *far_pairs)
{
    if (order == 2 && !t->has_quad)
        return WRAP_RAISE_ERROR(GRAV_FAILURE, "Quadrupole walk of a tree built without second moments");
    if (phi > 0.0 && order != 1)
        return WRAP_RAISE_ERROR(GRAV_FAILURE, "Local expansions are built for the monopole group walk only");
    if (phi > 0.0 && !t->has_quad)
        return WRAP_RAISE_ERROR(GRAV_FAILURE, "Local expansions need a tree with second moments");
    WalkArgs args;
    args.theta2 = theta * theta;
    args.eps2 = eps * eps;
    args.G = G;
    args.phi2 = phi > 0.0 ? phi * phi : 0.0;
    args.thetaf2 = theta_far > theta ? theta_far * theta_far : args.theta2;
    void *p;
    GS_TRY(slot(c, SLOT_MISC, 256, &p));
    // [0]: interactions (+ overflow flag), [1]: far (group, node) pairs
    unsigned long long *cnt = reinterpret_cast<unsigned long long *>(static_cast<char *>(p) + 64);
    GS_HIP(hipMemsetAsync(cnt, 0, 2 * sizeof(unsigned long long), c->stream));
    // groups: at most one per particle (they partition the particles)
    void *gp;
    GS_TRY(slot(c, SLOT_NODE_TMP, sizeof(int2) * static_cast<size_t>(t->n) + 256, &gp));
    int2 *groups = static_cast<int2 *>(gp);
    unsigned *ng = reinterpret_cast<unsigned *>(static_cast<char *>(p) + 128);
    // parent list and per-parent group counts, scanned into offsets (the
    // sort's scratch slots; the parent list in the groups' slot's tail)
    void *hp;
    GS_TRY(slot(c, SLOT_HIST, sizeof(unsigned) * static_cast<size_t>(t->num_nodes) + 256, &hp));
    unsigned *offs = static_cast<unsigned *>(hp);
    GS_TRY(slot(c, SLOT_HIST_SCAN, scan_temp_bytes(t->num_nodes) + 64, &hp));
    unsigned *stmp = static_cast<unsigned *>(hp);
    // (and after it, each parent's wide-run bits)
    GS_TRY(slot(c, SLOT_GROUP_PARENTS, 2 * sizeof(int) * static_cast<size_t>(t->num_nodes) + 256, &hp));
    int *parents = static_cast<int *>(hp);
    unsigned *wide = reinterpret_cast<unsigned *>(parents + t->num_nodes);
    const int rec = timing_begin(c, "bh_walk");
    const double cap = group_width_cap(t->width, t->n);
    const dim3 ngrid((t->num_nodes + 255) / 256);
    GS_HIP(scan_exclusive(c->stream, t->gflag, offs, t->num_nodes, stmp, ng));
    hipLaunchKernelGGL(k_group_parent_list, ngrid, dim3(256), 0, c->stream, t->num_nodes, offs, ng,
                       parents);
    GS_HIP(hipGetLastError());
    unsigned num_parents = 0;
    GS_HIP(hipMemcpyAsync(&num_parents, ng, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    if (num_parents > static_cast<unsigned>(t->num_nodes))
        return WRAP_RAISE_ERROR(GRAV_FAILURE, "Barnes-Hut group parents beyond the nodes");
    const dim3 pgrid((num_parents + 3) / 4);  // 4 wavefronts per workgroup
    hipLaunchKernelGGL(k_group_runs<false>, pgrid, dim3(256), 0, c->stream, t->nodes, parents,
                       static_cast<int>(num_parents), t->n, t->P, cap, offs, wide, groups);
    GS_HIP(hipGetLastError());
    GS_HIP(scan_exclusive(c->stream, offs, offs, num_parents, stmp, ng));
    unsigned num_groups = 0;
    GS_HIP(hipMemcpyAsync(&num_groups, ng, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    if (num_groups > static_cast<unsigned>(t->n))
        return WRAP_RAISE_ERROR(GRAV_FAILURE, "Barnes-Hut group list larger than the particles");
    hipLaunchKernelGGL(k_group_runs<true>, pgrid, dim3(256), 0, c->stream, t->nodes, parents,
                       static_cast<int>(num_parents), t->n, t->P, cap, offs, wide, groups);
    GS_HIP(hipGetLastError());
    const int grid = (static_cast<int>(num_groups) + GW_WAVES - 1) / GW_WAVES;
    if (order >= 2)
        hipLaunchKernelGGL((k_walk_group<true, false>), dim3(grid), dim3(GW_BLOCK), 0, c->stream,
                           t->nodes, t->P, t->sidx, t->n, groups, static_cast<int>(num_groups), args,
                           a, cnt, t->quadf);
    else if (args.phi2 > 0.0)
        hipLaunchKernelGGL((k_walk_group<false, true>), dim3(grid), dim3(GW_BLOCK), 0, c->stream,
                           t->nodes, t->P, t->sidx, t->n, groups, static_cast<int>(num_groups), args,
                           a, cnt, t->quadf);
    else
        hipLaunchKernelGGL((k_walk_group<false, false>), dim3(grid), dim3(GW_BLOCK), 0, c->stream,
                           t->nodes, t->P, t->sidx, t->n, groups, static_cast<int>(num_groups), args,
                           a, cnt, t->quadf);
    timing_end(c, rec);
    GS_HIP(hipGetLastError());
    unsigned long long h[2] = {0, 0};
    GS_HIP(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    if (h[0] >> 62) return WRAP_RAISE_ERROR(GRAV_FAILURE, "Barnes-Hut group walk stack overflow");
    if (interactions) *interactions = static_cast<long long>(h[0]);
    if (far_pairs) *far_pairs = static_cast<long long>(h[1]);
    return make_success_error_status();
}

ErrorStatus tree_to_host(Context *c, DeviceTree *t, LinearOctree *out)
{
    const int n = t->n, nn = t->num_nodes;
    std::vector<NodeRec> h(nn);
    std::vector<double4> hm(nn);
    GS_HIP(hipMemcpyAsync(h.data(), t->nodes, sizeof(NodeRec) * nn, hipMemcpyDeviceToHost,
                          c->stream));
    GS_HIP(hipMemcpyAsync(hm.data(), t->mom, sizeof(double4) * nn, hipMemcpyDeviceToHost,
                          c->stream));
    *out = get_new_linear_octree();
    out->box_width = t->width;
    out->num_internal_nodes = nn;
    out->particle_morton_indices_deepest_level = static_cast<int64 *>(malloc(sizeof(int64) * n));
    out->sorted_indices = static_cast<int *>(malloc(sizeof(int) * n));
    out->tree_num_particles = static_cast<int *>(malloc(sizeof(int) * nn));
    out->tree_num_internal_children = static_cast<int *>(malloc(sizeof(int) * nn));
    out->tree_first_particle_sorted_idx = static_cast<int *>(malloc(sizeof(int) * nn));
    out->tree_first_internal_children_idx = static_cast<int *>(malloc(sizeof(int) * nn));
    out->tree_mass = static_cast<double *>(malloc(sizeof(double) * nn));
    out->tree_center_of_mass_x = static_cast<double *>(malloc(sizeof(double) * nn));
    out->tree_center_of_mass_y = static_cast<double *>(malloc(sizeof(double) * nn));
    out->tree_center_of_mass_z = static_cast<double *>(malloc(sizeof(double) * nn));
    if (!out->particle_morton_indices_deepest_level || !out->sorted_indices ||
        !out->tree_num_particles || !out->tree_num_internal_children ||
        !out->tree_first_particle_sorted_idx || !out->tree_first_internal_children_idx ||
        !out->tree_mass || !out->tree_center_of_mass_x || !out->tree_center_of_mass_y ||
        !out->tree_center_of_mass_z) {
        GS_HIP(hipStreamSynchronize(c->stream));
        free_linear_octree(out);
        return WRAP_RAISE_ERROR(GRAV_MEMORY_ERROR, "Failed to allocate memory for the octree");
    }
    GS_HIP(hipMemcpyAsync(out->particle_morton_indices_deepest_level, t->skeys, sizeof(int64) * n,
                          hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipMemcpyAsync(out->sorted_indices, t->sidx, sizeof(int) * n, hipMemcpyDeviceToHost,
                          c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    for (int i = 0; i < nn; i++) {
        out->tree_num_particles[i] = h[i].count;
        out->tree_num_internal_children[i] = h[i].nchild;
        out->tree_first_particle_sorted_idx[i] = h[i].first;
        out->tree_first_internal_children_idx[i] = h[i].child0;
        out->tree_mass[i] = hm[i].x;
        out->tree_center_of_mass_x[i] = h[i].cx;
        out->tree_center_of_mass_y[i] = h[i].cy;
        out->tree_center_of_mass_z[i] = h[i].cz;
    }
    return make_success_error_status();
}

}  // namespace gs
