// scan.hip -- device-wide exclusive prefix sum of uint32 (three kernels:
// per-tile sums, one-block scan of the tile sums, per-tile scan + offset).
// Used by the radix sort (digit-major histograms) and the octree build
// (child counts per level).  In-place operation is allowed.
#include "gs_internal.h"
#include "scan.h"

namespace gs {

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;
constexpr int SCAN_TOP_ITEMS = 16;
constexpr int SCAN_MAX_TILES = SCAN_BLOCK * SCAN_TOP_ITEMS;

// Exclusive scan of one value per thread across a 256-thread block.
__device__ unsigned block_exclusive_scan(unsigned v, unsigned *total)
{
    __shared__ unsigned buf[SCAN_BLOCK];
    const int t = threadIdx.x;
    buf[t] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
        unsigned add = t >= off ? buf[t - off] : 0u;
        __syncthreads();
        buf[t] += add;
        __syncthreads();
    }
    const unsigned incl = buf[t];
    *total = buf[SCAN_BLOCK - 1];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_tilesum(const unsigned *__restrict__ in,
                                                             long long n,
                                                             unsigned *__restrict__ tsum)
{
    const long long base = static_cast<long long>(blockIdx.x) * SCAN_TILE;
    unsigned s = 0;
    for (int k = 0; k < SCAN_ITEMS; k++) {
        const long long q = base + k * SCAN_BLOCK + threadIdx.x;
        if (q < n) s += in[q];
    }
    unsigned total;
    block_exclusive_scan(s, &total);
    if (threadIdx.x == 0) tsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_top(unsigned *__restrict__ tsum, int ntiles,
                                                         unsigned *__restrict__ total_out)
{
    unsigned v[SCAN_TOP_ITEMS];
    unsigned s = 0;
    for (int k = 0; k < SCAN_TOP_ITEMS; k++) {
        const int q = threadIdx.x * SCAN_TOP_ITEMS + k;
        v[k] = q < ntiles ? tsum[q] : 0u;
        s += v[k];
    }
    unsigned total;
    unsigned run = block_exclusive_scan(s, &total);
    for (int k = 0; k < SCAN_TOP_ITEMS; k++) {
        const int q = threadIdx.x * SCAN_TOP_ITEMS + k;
        if (q < ntiles) tsum[q] = run;
        run += v[k];
    }
    if (threadIdx.x == 0 && total_out) *total_out = total;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_apply(const unsigned *__restrict__ in,
                                                           unsigned *__restrict__ out,
                                                           long long n,
                                                           const unsigned *__restrict__ tsum)
{
    // thread t owns the contiguous items [t*ITEMS, (t+1)*ITEMS) of the tile
    const long long base = static_cast<long long>(blockIdx.x) * SCAN_TILE +
                           static_cast<long long>(threadIdx.x) * SCAN_ITEMS;
    unsigned v[SCAN_ITEMS];
    unsigned s = 0;
    for (int k = 0; k < SCAN_ITEMS; k++) {
        const long long q = base + k;
        v[k] = q < n ? in[q] : 0u;
        s += v[k];
    }
    unsigned total;
    unsigned run = block_exclusive_scan(s, &total) + tsum[blockIdx.x];
    for (int k = 0; k < SCAN_ITEMS; k++) {
        const long long q = base + k;
        if (q < n) out[q] = run;
        run += v[k];
    }
}

// Beyond SCAN_MAX_TILES tiles the tile sums are scanned by the same three
// kernels one level up (their temporary storage follows this level's).
long long scan_max_elements() { return static_cast<long long>(SCAN_TILE) * 0x7fffffffLL; }

size_t scan_temp_bytes(long long n)
{
    const long long tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    size_t b = static_cast<size_t>(tiles + 1) * sizeof(unsigned);
    if (tiles > SCAN_MAX_TILES) b += scan_temp_bytes(tiles);
    return b;
}

hipError_t scan_exclusive(hipStream_t stream, const unsigned *in, unsigned *out, long long n,
                          unsigned *temp, unsigned *total_out)
{
    if (n <= 0) return hipSuccess;
    const long long tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (n > scan_max_elements()) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_scan_tilesum, dim3(static_cast<unsigned>(tiles)), dim3(SCAN_BLOCK), 0,
                       stream, in, n, temp);
    if (tiles <= SCAN_MAX_TILES) {
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_BLOCK), 0, stream, temp,
                           static_cast<int>(tiles), total_out);
    } else {
        const hipError_t e = scan_exclusive(stream, temp, temp, tiles, temp + tiles + 1, total_out);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_scan_apply, dim3(static_cast<unsigned>(tiles)), dim3(SCAN_BLOCK), 0,
                       stream, in, out, n, temp);
    return hipGetLastError();
}

}  // namespace gs
