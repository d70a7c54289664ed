// mlp.hip -- the learned Barnes-Hut correction: a 16 -> 128 -> 128 -> 128 -> 3
// MLP with tanh-GELU, trained and evaluated on the GPU with BF16 MFMA
// (v_mfma_f32_16x16x32_bf16) and FP32 accumulation.
//
//   k_gemm        C = op(A) op(B) (+ epilogues: bias, bias+GELU with the
//                 pre-activation kept for the backward pass; the product by
//                 GELU'(Z) in the backward pass).  One wavefront
//                 per 16x16 output tile, K in steps of 32; operands are read
//                 as FP32 (two float4 per fragment where k is contiguous) and
//                 rounded to BF16 (nearest even) into the MFMA fragments.
//                 The weight gradients (few tiles, K = batch rows) are split
//                 over K into slices that k_reduce_grads adds in a fixed
//                 order.  Fragment layout (16x16x32): lane l holds
//                 A[l%16][k0 + 8(l/16) + 0..7], B[k0 + 8(l/16) + 0..7][l%16];
//                 the result D[4(l/16) + 0..3][l%16].
//                 The last forward layer's epilogue in training forms the
//                 MSE gradient and the loss partials; the weight-gradient
//                 products also sum the bias gradients' columns.
//   k_reduce_grads  adds the gradient slices (fixed order), then Adam on the
//                 FP32 master weights
//   k_features    per-particle input features from the coarse walk
//   k_mlp_infer   fused inference + correction for method BARNES_HUT_LEARNED:
//                 a = a_coarse + G (q + o) + G s3 * tscale * (y in the local frame)
//                 (q, o: the walk's quadrupole and octupole sums, the physics
//                 part of the predicted residual; s3 = sum M l^3 / d^5, the
//                 scale of the next multipole orders)
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

#include "gs_internal.h"
#include "mlp.h"

namespace gs {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __bf16 to_bf16(float f)
{
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    const unsigned short h = static_cast<unsigned short>(u >> 16);
    return __builtin_bit_cast(__bf16, h);
}

// Eight consecutive bf16 (16 bytes, aligned) from LDS.
__device__ __forceinline__ bf16x8 lds8(const __bf16 *p)
{
    return *reinterpret_cast<const bf16x8 *>(p);
}

// GELU, tanh form, through the logistic function: (1 + tanh u) / 2 =
// 1 / (1 + exp(-2u)) and 1 - tanh^2 u = 4 sg (1 - sg), u = sqrt(2/pi) (z + 0.044715 z^3)
__device__ __forceinline__ float gelu_sg(float z)
{
    const float u = 0.7978845608028654f * (z + 0.044715f * z * z * z);
    return __builtin_amdgcn_rcpf(1.0f + __expf(-2.0f * u));
}

__device__ __forceinline__ float gelu(float z) { return z * gelu_sg(z); }

__device__ __forceinline__ float gelu_grad(float z)
{
    const float sg = gelu_sg(z);
    return sg + z * (2.0f * sg * (1.0f - sg)) * 0.7978845608028654f * (1.0f + 0.134145f * z * z);
}

// Eight consecutive k of one operand row into a BF16 fragment.  KC: the k
// index is contiguous and 16-byte aligned (two float4 loads); otherwise eight
// loads of stride sk.  Entries at k >= kend are zero.  SUM: also adds the
// eight FP32 values to *sum.
template <bool KC, bool SUM = false>
__device__ __forceinline__ bf16x8 load_frag(const float *__restrict__ row, long long sk, int kq,
                                            int kend, bool row_ok, float *sum = nullptr)
{
    bf16x8 x;
    float u[8];
    if (KC) {
        float4 u0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), u1 = u0;
        if (row_ok && kq < kend) {  // K % 8 == 0: the eight are all in range or all out
            const float4 *p = reinterpret_cast<const float4 *>(row + kq);
            u0 = p[0];
            u1 = p[1];
        }
        u[0] = u0.x; u[1] = u0.y; u[2] = u0.z; u[3] = u0.w;
        u[4] = u1.x; u[5] = u1.y; u[6] = u1.z; u[7] = u1.w;
    } else {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int k = kq + j;
            u[j] = (row_ok && k < kend) ? row[k * sk] : 0.0f;
        }
    }
#pragma unroll
    for (int j = 0; j < 8; j++) x[j] = to_bf16(u[j]);
    if (SUM) {
        float t = *sum;
#pragma unroll
        for (int j = 0; j < 8; j++) t += u[j];
        *sum = t;
    }
    return x;
}

// One wavefront's 16x16 tile over k in [kbeg, kend), 128 k at a time: the
// fragments of all four 32-k steps are loaded before their products are
// issued (slices and the hidden width are 128 k: one memory latency each).
template <bool A_KC, bool B_KC, bool SUM>
__device__ __forceinline__ f32x4 gemm_loop(const float *__restrict__ arow, long long sak, bool aok,
                                           const float *__restrict__ bcol, long long sbk, bool bok,
                                           int kbeg, int kend, int q, float *cs)
{
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = kbeg; k0 < kend; k0 += 128) {
        bf16x8 a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int kq = k0 + 32 * u + 8 * q;
            a[u] = load_frag<A_KC>(arow, sak, kq, kend, aok);
            b[u] = load_frag<B_KC, SUM>(bcol, sbk, kq, kend, bok, cs);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[u], b[u], acc, 0, 0, 0);
    }
    return acc;
}

// Epilogue operands of k_gemm.
struct GemmEpi {
    int mode;           // see k_gemm
    const float *bias;  // modes 1, 2, 4
    float *Z;           // mode 1: written; mode 3: read
    const float *T;     // mode 4: targets, rows x N
    float *loss;        // mode 4: one partial sum of (Y - T)^2 per wavefront
    float inv_rows;     // mode 4
    float *bsum;        // column sums of B over each slice of K (weight gradients), or null
};

// C(m, n) = sum_k A(m, k) B(k, n); A(m, k) = A[m * sam + k * sak], B likewise.
// mode 0: C = acc; 1: Z = acc + bias, C = GELU(Z); 2: C = acc + bias;
// 3: C = acc GELU'(Z) (the backward pass through the activation);
// 4: Y = acc + bias, C = 2 (Y - T) / rows, loss partials (the MSE gradient).
// Split K: workgroup z sums k in [z kchunk, (z + 1) kchunk) into C + z csplit,
// and, with bsum, the column sums of B over the same k into bsum + z csplit
// (the bias gradient's slice).  A_KC / B_KC: sak == 1 / sbk == 1 with rows
// 16-byte aligned and K % 8 == 0.
template <bool A_KC, bool B_KC>
__global__ __launch_bounds__(256) void k_gemm(const float *__restrict__ A, long long sam,
                                              long long sak, const float *__restrict__ B,
                                              long long sbk, long long sbn, int M, int N, int K,
                                              int kchunk, float *__restrict__ C, long long ldc,
                                              long long csplit, GemmEpi ep)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int tm = (blockIdx.x * 2 + (wave >> 1)) * 16;
    const int tn = (blockIdx.y * 2 + (wave & 1)) * 16;
    const int slot = (blockIdx.x * 2 + (wave >> 1)) * gridDim.y * 2 + blockIdx.y * 2 + (wave & 1);
    if (tm >= M || tn >= N) {
        if (ep.mode == 4 && lane == 0) ep.loss[slot] = 0.0f;
        return;
    }
    const int r = lane & 15, q = lane >> 4;
    const int kbeg = blockIdx.z * kchunk;
    const int kend = min(K, kbeg + kchunk);
    C += blockIdx.z * csplit;
    const int am = tm + r, bn = tn + r;
    const float *arow = A + static_cast<long long>(min(am, M - 1)) * sam;
    const float *bcol = B + static_cast<long long>(min(bn, N - 1)) * sbn;
    f32x4 acc;
    if (ep.bsum && tm == 0) {
        // the first row of tiles also sums its B columns (FP32, fixed order)
        float cs = 0.0f;
        acc = gemm_loop<A_KC, B_KC, true>(arow, sak, am < M, bcol, sbk, bn < N, kbeg, kend, q, &cs);
        cs += __shfl_xor(cs, 16);
        cs += __shfl_xor(cs, 32);
        if (q == 0 && bn < N) ep.bsum[blockIdx.z * csplit + bn] = cs;
    } else {
        acc = gemm_loop<A_KC, B_KC, false>(arow, sak, am < M, bcol, sbk, bn < N, kbeg, kend, q, nullptr);
    }
    const int cn = tn + r;
    float l = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int cm = tm + 4 * q + j;
        if (cm >= M || cn >= N) continue;
        float v = acc[j];
        const int mode = ep.mode;
        if (mode == 1 || mode == 2 || mode == 4) v += ep.bias[cn];
        if (mode == 1) {
            ep.Z[cm * ldc + cn] = v;
            v = gelu(v);
        } else if (mode == 3) {
            v *= gelu_grad(ep.Z[cm * ldc + cn]);
        } else if (mode == 4) {
            const float d = v - ep.T[cm * ldc + cn];
            l += d * d;
            v = 2.0f * d * ep.inv_rows;
        }
        C[cm * ldc + cn] = v;
    }
    if (ep.mode == 4) {
        // wavefront sum in a fixed order
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
        if (lane == 0) ep.loss[slot] = l;
    }
}

// Number of loss partials k_gemm writes in mode 4 for M x N.
static int gemm_loss_parts(int M, int N) { return ((M + 31) / 32) * 2 * ((N + 31) / 32) * 2; }

// Split-K slices of the weight gradients: every weight-gradient product has
// K = batch rows; at most MLP_SPLIT_MAX slices of at least 128 rows, each a
// multiple of 32 rows but the last.
constexpr int MLP_SPLIT_MAX = 32;
static void wgrad_slices(int rows, int *splits, int *kchunk)
{
    const int want = std::max(1, std::min(MLP_SPLIT_MAX, rows / 128));
    const int kc = ((rows + want - 1) / want + 31) / 32 * 32;
    *kchunk = kc;
    *splits = std::max(1, (rows + kc - 1) / kc);  // <= want
}

static bool kc_ok(const float *p, long long s_row, long long s_k, int K)
{
    return s_k == 1 && s_row % 4 == 0 && K % 8 == 0 && (reinterpret_cast<uintptr_t>(p) & 15) == 0;
}

static void launch_gemm(Context *c, dim3 grid, const float *A, long long sam, long long sak,
                        const float *B, long long sbk, long long sbn, int M, int N, int K,
                        int kchunk, float *C, long long ldc, long long csplit, const GemmEpi &ep)
{
    const bool akc = kc_ok(A, sam, sak, K), bkc = kc_ok(B, sbn, sbk, K);
    auto launch = [&](auto kern) {
        hipLaunchKernelGGL(kern, grid, dim3(256), 0, c->stream, A, sam, sak, B, sbk, sbn, M, N, K,
                           kchunk, C, ldc, csplit, ep);
    };
    if (akc && bkc) launch(k_gemm<true, true>);
    else if (akc) launch(k_gemm<true, false>);
    else if (bkc) launch(k_gemm<false, true>);
    else launch(k_gemm<false, false>);
}

// C = op(A) op(B) with the epilogue of ep (no split).
static hipError_t gemm(Context *c, const float *A, long long sam, long long sak, const float *B,
                       long long sbk, long long sbn, int M, int N, int K, float *C, long long ldc,
                       const GemmEpi &ep)
{
    dim3 grid((M + 31) / 32, (N + 31) / 32, 1);
    launch_gemm(c, grid, A, sam, sak, B, sbk, sbn, M, N, K, K, C, ldc, 0, ep);
    return hipGetLastError();
}

// Weight gradient dW = A^T B (A: K x M activations, B: K x N output
// gradients, both row-major with K = rows) and the bias gradient's column
// sums of B, as `splits` slices over K: slice z into P + z MLP_NPARAM at
// the weight's and the bias's parameter offsets (k_reduce_grads adds them).
static hipError_t gemm_wgrad(Context *c, const float *A, int M, const float *B, int N, int rows,
                             float *P, int off_w, int off_b, int splits, int kchunk)
{
    if (splits < 1 || splits > MLP_SPLIT_MAX || static_cast<long long>(splits) * kchunk < rows ||
        static_cast<long long>(splits - 1) * kchunk >= rows)
        return hipErrorInvalidValue;
    GemmEpi ep{};
    ep.bsum = P + off_b;
    dim3 grid((M + 31) / 32, (N + 31) / 32, splits);
    launch_gemm(c, grid, A, 1, M, B, N, 1, M, N, rows, kchunk, P + off_w, N, MLP_NPARAM, ep);
    return hipGetLastError();
}


// ------------------------------------------------------------------ fused training step
// k_train_slice: forward and backward of one 128-row slice of a batch in one
// workgroup (8 wavefronts, one 16-row tile each), the gradient slices into
// P + z MLP_NPARAM as gemm_wgrad writes them (k_reduce_grads adds them).
// The products, their BF16 operands and FP32 accumulation over k in steps of
// 32 are those of k_gemm; the forward layers of a wavefront only read its
// own rows, so only the weight gradients (sums over all rows of the slice)
// wait at a workgroup barrier.  Weights W2, W3 are resident in LDS as BF16
// in both layouts the products read (outputs x inputs for the forward pass,
// inputs x outputs for the backward pass), W1 and W4 transposed, copied from
// a weight image in device memory that the Adam step keeps up to date; the
// activations go through global memory (L2): GELU' at the pre-activations
// (FP32, formed in the forward pass with the same tanh as GELU) and BF16
// activations and output gradients in row and transposed layouts, so that
// every fragment is one 16-byte read.
constexpr int TS_ROWS = 128;              // rows per slice (= per workgroup)
constexpr int TS_WAVES = TS_ROWS / 16;    // one 16-row tile per wavefront
constexpr int TS_LD = MLP_H + 8;          // LDS row of 128 bf16, padded by 16 B
constexpr int TS_LD1 = MLP_F + 8;         // W1^T rows: K = 16, padded by 16 B
constexpr int TS_MAX_SLICES = 32;         // batches of up to 4096 rows
static_assert(TS_MAX_SLICES <= MLP_SPLIT_MAX, "the gradient slices fit the split buffer");
constexpr size_t TS_HH = static_cast<size_t>(TS_ROWS) * MLP_H;  // one 128 x 128 block

struct alignas(16) TsLds {
    __bf16 w1t[MLP_H * TS_LD1];
    __bf16 w2t[MLP_H * TS_LD], w3t[MLP_H * TS_LD];  // [out][in]
    __bf16 w2n[MLP_H * TS_LD], w3n[MLP_H * TS_LD];  // [in][out]
    __bf16 w4t[16 * TS_LD];                         // [out][in], outputs 3..15 zero
    float b1[MLP_H], b2[MLP_H], b3[MLP_H], b4[4];
    float cs[2][TS_WAVES][MLP_H];  // per-wavefront column sums of dZ (two layers in flight)
    float cs4[TS_WAVES][4];        // per-wavefront column sums of dY
};

// The weights as k_train_slice holds them in LDS: the first TS_IMG_BYTES of
// TsLds, kept in device memory (k_weight_image writes it whole; the Adam
// step writes the entries of the parameters it updates), so that a
// workgroup fills its LDS with 16-byte copies.
constexpr size_t TS_IMG_BYTES = offsetof(TsLds, cs);
static_assert(TS_IMG_BYTES % 16 == 0, "the weight image is copied in 16-byte pieces");
constexpr int TS_I_W1T = offsetof(TsLds, w1t) / 2, TS_I_W2T = offsetof(TsLds, w2t) / 2,
              TS_I_W3T = offsetof(TsLds, w3t) / 2, TS_I_W2N = offsetof(TsLds, w2n) / 2,
              TS_I_W3N = offsetof(TsLds, w3n) / 2, TS_I_W4T = offsetof(TsLds, w4t) / 2;  // bf16 units
constexpr int TS_F_B1 = offsetof(TsLds, b1) / 4, TS_F_B2 = offsetof(TsLds, b2) / 4,
              TS_F_B3 = offsetof(TsLds, b3) / 4, TS_F_B4 = offsetof(TsLds, b4) / 4;  // float units

// Writes parameter i (value v) into the weight image.
__device__ __forceinline__ void ts_image_put(char *img, int i, float v)
{
    __bf16 *h = reinterpret_cast<__bf16 *>(img);
    float *f = reinterpret_cast<float *>(img);
    if (i < MLP_OFF_B1) {
        const int in = (i - MLP_OFF_W1) / MLP_H, out = (i - MLP_OFF_W1) % MLP_H;
        h[TS_I_W1T + out * TS_LD1 + in] = to_bf16(v);
    } else if (i < MLP_OFF_W2) {
        f[TS_F_B1 + i - MLP_OFF_B1] = v;
    } else if (i < MLP_OFF_B2) {
        const int in = (i - MLP_OFF_W2) / MLP_H, out = (i - MLP_OFF_W2) % MLP_H;
        h[TS_I_W2T + out * TS_LD + in] = to_bf16(v);
        h[TS_I_W2N + in * TS_LD + out] = to_bf16(v);
    } else if (i < MLP_OFF_W3) {
        f[TS_F_B2 + i - MLP_OFF_B2] = v;
    } else if (i < MLP_OFF_B3) {
        const int in = (i - MLP_OFF_W3) / MLP_H, out = (i - MLP_OFF_W3) % MLP_H;
        h[TS_I_W3T + out * TS_LD + in] = to_bf16(v);
        h[TS_I_W3N + in * TS_LD + out] = to_bf16(v);
    } else if (i < MLP_OFF_W4) {
        f[TS_F_B3 + i - MLP_OFF_B3] = v;
    } else if (i < MLP_OFF_B4) {
        const int in = (i - MLP_OFF_W4) / MLP_O, out = (i - MLP_OFF_W4) % MLP_O;
        h[TS_I_W4T + out * TS_LD + in] = to_bf16(v);
    } else {
        f[TS_F_B4 + i - MLP_OFF_B4] = v;
    }
}

// Every parameter of w into the weight image (the padding, among it the
// unused rows 3..15 of W4^T that are read as zero weights, is cleared
// before).
__global__ void k_weight_image(const float *__restrict__ w, char *__restrict__ img)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < MLP_NPARAM) ts_image_put(img, t, w[t]);
}

// Global scratch of one slice.
struct TsSlice {
    float *Z;     // [3][128][128] GELU' at the pre-activations
    __bf16 *H;    // [3][2][128 x 128] activations: [row][col], [col][row]
    __bf16 *D;    // [2][2][128 x 128] output gradients (two pairs): [row][col], [col][row]
    __bf16 *Xt;   // [16][128] inputs, transposed
    float *dY;    // [128][3]
};
constexpr size_t TS_SLICE_BYTES = sizeof(float) * 3 * TS_HH + sizeof(__bf16) * 6 * TS_HH +
                                  sizeof(__bf16) * 4 * TS_HH + sizeof(__bf16) * MLP_F * TS_ROWS +
                                  sizeof(float) * TS_ROWS * 4;

__device__ __forceinline__ TsSlice ts_slice(char *base, int z)
{
    char *p = base + static_cast<size_t>(z) * TS_SLICE_BYTES;
    TsSlice t;
    t.Z = reinterpret_cast<float *>(p);
    p += sizeof(float) * 3 * TS_HH;
    t.H = reinterpret_cast<__bf16 *>(p);
    p += sizeof(__bf16) * 6 * TS_HH;
    t.D = reinterpret_cast<__bf16 *>(p);
    p += sizeof(__bf16) * 4 * TS_HH;
    t.Xt = reinterpret_cast<__bf16 *>(p);
    p += sizeof(__bf16) * MLP_F * TS_ROWS;
    t.dY = reinterpret_cast<float *>(p);
    return t;
}

__device__ __forceinline__ bf16x8 gld8(const __bf16 *p)
{
    return *reinterpret_cast<const bf16x8 *>(p);
}

// Epilogue of a forward hidden layer for one 16x16 tile (rows tm + 4q + j,
// column col): H = GELU(acc + b) in both layouts, and GELU'(acc + b) (FP32)
// for the backward pass (one exponential for both).
__device__ __forceinline__ void ts_store_hidden(const f32x4 &acc, float bias, int tm, int q, int col,
                                                float *Gp, __bf16 *Hr, __bf16 *Ht)
{
    __bf16 h4[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int row = tm + 4 * q + j;
        const float z = acc[j] + bias;
        const float sg = gelu_sg(z);  // gelu() and gelu_grad() from one exponential
        Gp[row * MLP_H + col] = sg + z * (2.0f * sg * (1.0f - sg)) * 0.7978845608028654f * (1.0f + 0.134145f * z * z);
        h4[j] = to_bf16(z * sg);
        Hr[row * MLP_H + col] = h4[j];
    }
    // four consecutive rows of one column: one 8-byte store
    typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
    bf16x4 v4 = {h4[0], h4[1], h4[2], h4[3]};
    *reinterpret_cast<bf16x4 *>(Ht + col * TS_ROWS + tm + 4 * q) = v4;
}

// GELU' of this lane's 4 rows of the 8 column tiles (loaded before the products).
__device__ __forceinline__ void ts_load_gp(const float *Gp, int tm, int q, int r, float gp[8][4])
{
#pragma unroll
    for (int nt = 0; nt < MLP_H / 16; nt++)
#pragma unroll
        for (int j = 0; j < 4; j++) gp[nt][j] = Gp[(tm + 4 * q + j) * MLP_H + nt * 16 + r];
}

// Epilogue of a backward product through GELU for one tile: dZ = acc GELU'
// in both layouts, and this lane's sum over its four rows (FP32).
__device__ __forceinline__ float ts_store_dz(const f32x4 &acc, int tm, int q, int col, const float gp[4],
                                             __bf16 *Dr, __bf16 *Dt)
{
    __bf16 d4[4];
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int row = tm + 4 * q + j;
        const float v = acc[j] * gp[j];
        s += v;
        d4[j] = to_bf16(v);
        Dr[row * MLP_H + col] = d4[j];
    }
    typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
    bf16x4 v4 = {d4[0], d4[1], d4[2], d4[3]};
    *reinterpret_cast<bf16x4 *>(Dt + col * TS_ROWS + tm + 4 * q) = v4;
    return s;
}

// Sum over the four row groups q of a lane's value (lanes r, r+16, r+32, r+48).
__device__ __forceinline__ float ts_sum_q(float v)
{
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
}

// dW tiles (input tiles it0, it0 + 1; output tiles ot0 .. ot0 + 3) = A^T B
// over the slice's 128 rows, with A^T and B both transposed ([col][row],
// 16-byte fragments; each fragment is read once per wavefront).
__device__ __forceinline__ void ts_wgrad_block(const __bf16 *At, const __bf16 *Bt, int it0, int ot0,
                                               int r, int q, f32x4 acc[2][4])
{
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int o = 0; o < 4; o++) acc[i][o] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ks = 0; ks < 4; ks++) {
        bf16x8 a[2], b[4];
#pragma unroll
        for (int i = 0; i < 2; i++) a[i] = gld8(At + ((it0 + i) * 16 + r) * TS_ROWS + ks * 32 + 8 * q);
#pragma unroll
        for (int o = 0; o < 4; o++) b[o] = gld8(Bt + ((ot0 + o) * 16 + r) * TS_ROWS + ks * 32 + 8 * q);
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int o = 0; o < 4; o++)
                acc[i][o] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[o], acc[i][o], 0, 0, 0);
    }
}

// One dW tile (input tile it, output tile ot), as ts_wgrad_block.
__device__ __forceinline__ f32x4 ts_wgrad_tile(const __bf16 *At, const __bf16 *Bt, int it, int ot,
                                               int r, int q)
{
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    bf16x8 a[4], b[4];
#pragma unroll
    for (int ks = 0; ks < 4; ks++) {
        a[ks] = gld8(At + (it * 16 + r) * TS_ROWS + ks * 32 + 8 * q);
        b[ks] = gld8(Bt + (ot * 16 + r) * TS_ROWS + ks * 32 + 8 * q);
    }
#pragma unroll
    for (int ks = 0; ks < 4; ks++) acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ks], b[ks], acc, 0, 0, 0);
    return acc;
}

// Stores a weight-gradient tile (rows: inputs it*16 + 4q + j, column: output
// ot*16 + r of an (nin x nout) matrix) into the slice P.
__device__ __forceinline__ void ts_put_wgrad(float *P, int off, int nout, int it, int ot, int r,
                                             int q, const f32x4 &acc)
{
    const int o = ot * 16 + r;
    if (o >= nout) return;
#pragma unroll
    for (int j = 0; j < 4; j++) P[off + (it * 16 + 4 * q + j) * nout + o] = acc[j];
}

__global__ __launch_bounds__(64 * TS_WAVES) void k_train_slice(
    const float *__restrict__ X, const float *__restrict__ T, const int *__restrict__ idx, int rows,
    float inv_rows, const char *__restrict__ img, float *__restrict__ P, float *__restrict__ loss,
    char *__restrict__ scratch)
{
    __shared__ TsLds S;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int r = lane & 15, q = lane >> 4;
    const int z = blockIdx.x;
    const int r0 = z * TS_ROWS;
    const int nr = min(TS_ROWS, rows - r0);  // valid rows of this slice (>= 1)
    const TsSlice sl = ts_slice(scratch, z);
    float *Pz = P + static_cast<size_t>(z) * MLP_NPARAM;

    // weights into LDS: 16-byte copies of the image
    {
        const uint4 *src4 = reinterpret_cast<const uint4 *>(img);
        uint4 *dst4 = reinterpret_cast<uint4 *>(&S);
        constexpr int N16 = static_cast<int>(TS_IMG_BYTES / 16);
#pragma unroll 8
        for (int i = tid; i < N16; i += 64 * TS_WAVES) dst4[i] = src4[i];
    }
    __syncthreads();

    const int tm = wave * 16;  // this wavefront's rows within the slice
    const int row = tm + r;    // the A-fragment row of this lane
    const bool row_ok = row < nr;
    const long long src = row_ok ? (idx ? idx[r0 + row] : r0 + row) : 0;
    __bf16 *H1r = sl.H, *H1t = sl.H + TS_HH, *H2r = sl.H + 2 * TS_HH, *H2t = sl.H + 3 * TS_HH;
    __bf16 *H3r = sl.H + 4 * TS_HH, *H3t = sl.H + 5 * TS_HH;
    float *Z1 = sl.Z, *Z2 = sl.Z + TS_HH, *Z3 = sl.Z + 2 * TS_HH;

    // ---- forward, layer 1 (K = 16: one step with the upper half of k zero)
    {
        bf16x8 a;
        float u[8];
#pragma unroll
        for (int j = 0; j < 8; j++) u[j] = 0.0f;
        if (q < 2 && row_ok) {
            const float4 *xp = reinterpret_cast<const float4 *>(X + src * MLP_F + 8 * q);
            const float4 u0 = xp[0], u1 = xp[1];
            u[0] = u0.x; u[1] = u0.y; u[2] = u0.z; u[3] = u0.w;
            u[4] = u1.x; u[5] = u1.y; u[6] = u1.z; u[7] = u1.w;
        }
#pragma unroll
        for (int j = 0; j < 8; j++) a[j] = to_bf16(u[j]);
        if (q < 2) {
#pragma unroll
            for (int j = 0; j < 8; j++) sl.Xt[(8 * q + j) * TS_ROWS + row] = a[j];
        }
        const bf16x8 zero8 = {};
#pragma unroll
        for (int nt = 0; nt < MLP_H / 16; nt++) {
            const bf16x8 b = q < 2 ? lds8(S.w1t + (nt * 16 + r) * TS_LD1 + 8 * q) : zero8;
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
            const int col = nt * 16 + r;
            ts_store_hidden(acc, S.b1[col], tm, q, col, Z1, H1r, H1t);
        }
    }
    __syncthreads();
    // ---- forward, layers 2 and 3 (this wavefront's rows only)
#pragma unroll 1
    for (int layer = 2; layer <= 3; layer++) {
        const __bf16 *Hin = layer == 2 ? H1r : H2r;
        const __bf16 *Wt = layer == 2 ? S.w2t : S.w3t;
        const float *bias = layer == 2 ? S.b2 : S.b3;
        float *Zo = layer == 2 ? Z2 : Z3;
        __bf16 *Hor = layer == 2 ? H2r : H3r, *Hot = layer == 2 ? H2t : H3t;
        bf16x8 a[4];
#pragma unroll
        for (int ks = 0; ks < 4; ks++) a[ks] = gld8(Hin + row * MLP_H + ks * 32 + 8 * q);
#pragma unroll
        for (int nt = 0; nt < MLP_H / 16; nt++) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ks = 0; ks < 4; ks++)
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    a[ks], lds8(Wt + (nt * 16 + r) * TS_LD + ks * 32 + 8 * q), acc, 0, 0, 0);
            const int col = nt * 16 + r;
            ts_store_hidden(acc, bias[col], tm, q, col, Zo, Hor, Hot);
        }
        __syncthreads();
    }
    // ---- forward, layer 4 and the MSE gradient dY = 2 (Y - T) / rows
    {
        bf16x8 a[4];
#pragma unroll
        for (int ks = 0; ks < 4; ks++) a[ks] = gld8(H3r + row * MLP_H + ks * 32 + 8 * q);
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 4; ks++)
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ks], lds8(S.w4t + r * TS_LD + ks * 32 + 8 * q),
                                                          acc, 0, 0, 0);
        float l = 0.0f, cs = 0.0f;
        if (r < MLP_O) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int rr = tm + 4 * q + j;
                float dy = 0.0f;
                if (rr < nr) {
                    const long long s2 = idx ? idx[r0 + rr] : r0 + rr;
                    const float d = (acc[j] + S.b4[r]) - T[s2 * MLP_O + r];
                    l += d * d;
                    dy = 2.0f * d * inv_rows;
                }
                sl.dY[rr * MLP_O + r] = dy;
                cs += dy;
            }
        }
        // loss: wavefront sum in a fixed order
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
        if (lane == 0) loss[z * TS_WAVES + wave] = l;
        cs = ts_sum_q(cs);
        if (q == 0 && r < 4) S.cs4[wave][r] = cs;
    }
    __syncthreads();
    // ---- backward, layer 4: dZ3 = (dY W4^T) GELU'(Z3) (own rows) into pair 0
    __bf16 *D0r = sl.D, *D0t = sl.D + TS_HH, *D1r = sl.D + 2 * TS_HH, *D1t = sl.D + 3 * TS_HH;
    {
        float gp[8][4];
        ts_load_gp(Z3, tm, q, r, gp);
        bf16x8 a;
        float u[8];
#pragma unroll
        for (int j = 0; j < 8; j++) u[j] = (q == 0 && j < MLP_O) ? sl.dY[row * MLP_O + j] : 0.0f;
#pragma unroll
        for (int j = 0; j < 8; j++) a[j] = to_bf16(u[j]);
#pragma unroll
        for (int nt = 0; nt < MLP_H / 16; nt++) {
            // B(k = output, n = input) = W4[input][output], k < 3
            bf16x8 b;
            const int in = nt * 16 + r;
#pragma unroll
            for (int j = 0; j < 8; j++)
                b[j] = (q == 0 && j < MLP_O) ? S.w4t[j * TS_LD + in] : to_bf16(0.0f);
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
            const float s = ts_store_dz(acc, tm, q, in, gp[nt], D0r, D0t);
            const float cs = ts_sum_q(s);
            if (q == 0) S.cs[0][wave][in] = cs;
        }
    }
    __syncthreads();
    // dW4 = H3^T dY (input tile = wavefront, one output tile), db4, db3
    {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const bf16x8 a = gld8(H3t + (wave * 16 + r) * TS_ROWS + ks * 32 + 8 * q);
            bf16x8 b;
#pragma unroll
            for (int j = 0; j < 8; j++)
                b[j] = to_bf16(r < MLP_O ? sl.dY[(ks * 32 + 8 * q + j) * MLP_O + r] : 0.0f);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
        ts_put_wgrad(Pz, MLP_OFF_W4, MLP_O, wave, 0, r, q, acc);
        if (tid < MLP_O) {
            float s = 0.0f;
            for (int wv = 0; wv < TS_WAVES; wv++) s += S.cs4[wv][tid];
            Pz[MLP_OFF_B4 + tid] = s;
        }
        if (tid < MLP_H) {
            float s = 0.0f;
            for (int wv = 0; wv < TS_WAVES; wv++) s += S.cs[0][wv][tid];
            Pz[MLP_OFF_B3 + tid] = s;
        }
    }
    // ---- backward, layers 3 and 2: dW = H^T dZ (eight tiles per wavefront),
    // then dZ_below = (dZ W^T) GELU'(Z_below) (own rows)
#pragma unroll 1
    for (int layer = 3; layer >= 2; layer--) {
        const int cur = layer == 3 ? 0 : 1;  // cs buffer and D pair of this layer's dZ
        const __bf16 *Dr = cur == 0 ? D0r : D1r, *Dt = cur == 0 ? D0t : D1t;
        __bf16 *Nr = cur == 0 ? D1r : D0r, *Nt = cur == 0 ? D1t : D0t;
        const __bf16 *Hbt = layer == 3 ? H2t : H1t;
        const __bf16 *Wn = layer == 3 ? S.w3n : S.w2n;
        const float *Zb = layer == 3 ? Z2 : Z1;
        const int offw = layer == 3 ? MLP_OFF_W3 : MLP_OFF_W2;
        const int offb_below = layer == 3 ? MLP_OFF_B2 : MLP_OFF_B1;
        {
            // wavefront w: input tiles 2 (w / 2) + {0, 1}, output tiles 4 (w % 2) + {0..3}
            const int it0 = (wave >> 1) * 2, ot0 = (wave & 1) * 4;
            f32x4 acc[2][4];
            ts_wgrad_block(Hbt, Dt, it0, ot0, r, q, acc);
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int o = 0; o < 4; o++) ts_put_wgrad(Pz, offw, MLP_H, it0 + i, ot0 + o, r, q, acc[i][o]);
        }
        float gp[8][4];
        ts_load_gp(Zb, tm, q, r, gp);
        bf16x8 a[4];
#pragma unroll
        for (int ks = 0; ks < 4; ks++) a[ks] = gld8(Dr + row * MLP_H + ks * 32 + 8 * q);
#pragma unroll
        for (int nt = 0; nt < MLP_H / 16; nt++) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ks = 0; ks < 4; ks++)
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    a[ks], lds8(Wn + (nt * 16 + r) * TS_LD + ks * 32 + 8 * q), acc, 0, 0, 0);
            const int in = nt * 16 + r;
            const float s = ts_store_dz(acc, tm, q, in, gp[nt], Nr, Nt);
            const float cs = ts_sum_q(s);
            if (q == 0) S.cs[1 - cur][wave][in] = cs;
        }
        __syncthreads();
        if (tid < MLP_H) {
            float s = 0.0f;
            for (int wv = 0; wv < TS_WAVES; wv++) s += S.cs[1 - cur][wv][tid];
            Pz[offb_below + tid] = s;
        }
    }
    // ---- backward, layer 1: dW1 = X^T dZ1 (one input tile; output tile = wavefront)
    {
        f32x4 acc = ts_wgrad_tile(sl.Xt, D0t, 0, wave, r, q);
        ts_put_wgrad(Pz, MLP_OFF_W1, MLP_H, 0, wave, r, q, acc);
    }
}

// g[i] = sum over z < splits of P[z MLP_NPARAM + i] (fixed order); with
// adam, the Adam update of the FP32 master weights from g instead (bc1,
// bc2: the bias corrections 1 - b1^t, 1 - b2^t of step t), and with img the
// updated weight's entries of the weight image.
__global__ void k_reduce_grads(const float *__restrict__ P, int splits, float *__restrict__ g,
                               int adam, float *__restrict__ w, float *__restrict__ m,
                               float *__restrict__ v, float lr, float b1, float b2, float bc1,
                               float bc2, float eps, char *__restrict__ img)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= MLP_NPARAM) return;
    float gi = 0.0f;
#pragma unroll 8
    for (int z = 0; z < splits; z++) gi += P[static_cast<size_t>(z) * MLP_NPARAM + i];
    if (!adam) {
        g[i] = gi;
        return;
    }
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    w[i] = wi;
    if (img) ts_image_put(img, i, wi);
}

// Gradient slices of one batch of rows <= TS_MAX_SLICES * TS_ROWS through
// k_train_slice; returns the number of slices.  X, T: the rows (or, with
// idx, the rows idx[b]).
static ErrorStatus train_slices(Context *c, const float *X, const float *T, const int *idx, int rows,
                                const char *img, float *P, float *loss, int *splits)
{
    const int S = (rows + TS_ROWS - 1) / TS_ROWS;
    if (rows <= 0 || S > TS_MAX_SLICES)
        return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR, "fused training step: 1 to 4096 rows");
    void *scratch;
    GS_TRY(slot(c, SLOT_MLP_TS, TS_SLICE_BYTES * S, &scratch));
    hipLaunchKernelGGL(k_train_slice, dim3(S), dim3(64 * TS_WAVES), 0, c->stream, X, T, idx, rows,
                       1.0f / static_cast<float>(rows), img, P, loss, static_cast<char *>(scratch));
    GS_HIP(hipGetLastError());
    *splits = S;
    return make_success_error_status();
}

static bool fused_rows(int rows) { return rows >= 1 && rows <= TS_MAX_SLICES * TS_ROWS; }

// The weight image of w.
static ErrorStatus weight_image(Context *c, const float *w, char **img)
{
    void *p;
    GS_TRY(slot(c, SLOT_MLP_IMG, TS_IMG_BYTES, &p));
    GS_HIP(hipMemsetAsync(p, 0, TS_IMG_BYTES, c->stream));
    hipLaunchKernelGGL(k_weight_image, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, c->stream, w,
                       static_cast<char *>(p));
    GS_HIP(hipGetLastError());
    *img = static_cast<char *>(p);
    return make_success_error_status();
}

// Gather rows idx[b] of X (width F) into Xb (batch order).
__global__ void k_gather_rows(const float *__restrict__ X, const int *__restrict__ idx, int rows,
                              int F, float *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * F) return;
    const int r = i / F, f = i % F;
    out[i] = X[static_cast<long long>(idx[r]) * F + f];
}

// A training batch: rows idx[b] of the features X (MLP_F wide) and of the
// targets T (MLP_O wide), in batch order.
__global__ void k_gather_batch(const float *__restrict__ X, const float *__restrict__ T,
                               const int *__restrict__ idx, int rows, float *__restrict__ Xb,
                               float *__restrict__ Tb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    constexpr int W = MLP_F + MLP_O;
    if (i >= rows * W) return;
    const int r = i / W, f = i % W;
    const long long src = idx[r];
    if (f < MLP_F)
        Xb[r * MLP_F + f] = X[src * MLP_F + f];
    else
        Tb[r * MLP_O + f - MLP_F] = T[src * MLP_O + f - MLP_F];
}

hipError_t mlp_gather_rows(Context *c, const float *X, const int *idx, int rows, int F, float *out)
{
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_rows, dim3((rows * F + 255) / 256), dim3(256), 0, c->stream, X,
                       idx, rows, F, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------ model
static MlpModel g_model;

MlpModel &mlp_model() { return g_model; }

ErrorStatus mlp_ensure(Context *c)
{
    MlpModel &M = g_model;
    if (M.w) return make_success_error_status();
    GS_HIP(hipMalloc(&M.w, sizeof(float) * MLP_NPARAM));
    GS_HIP(hipMalloc(&M.g, sizeof(float) * MLP_NPARAM));
    GS_HIP(hipMalloc(&M.m, sizeof(float) * MLP_NPARAM));
    GS_HIP(hipMalloc(&M.v, sizeof(float) * MLP_NPARAM));
    GS_HIP(hipMemsetAsync(M.w, 0, sizeof(float) * MLP_NPARAM, c->stream));
    for (int k = 0; k < MLP_F; k++) {
        M.fmean[k] = 0.0f;
        M.fstd[k] = 1.0f;
    }
    M.tscale = 1.0f;
    return make_success_error_status();
}

// activations for `rows` rows: X, Z1..Z3, H1..H3, Y and their gradients
struct Acts {
    float *X, *Z1, *H1, *Z2, *H2, *Z3, *H3, *Y, *dY, *dH3, *dH2, *dH1, *loss;
    float *split;  // split-K slices of the gradients, MLP_SPLIT_MAX x MLP_NPARAM
};

static ErrorStatus acts(Context *c, int rows, Acts *a)
{
    const size_t H = static_cast<size_t>(rows) * MLP_H;
    const size_t nsplit = static_cast<size_t>(MLP_SPLIT_MAX) * MLP_NPARAM;
    const size_t need = sizeof(float) * (nsplit + static_cast<size_t>(rows) * MLP_F +
                                         9 * H + 2 * static_cast<size_t>(rows) * MLP_O + rows + 1024);
    void *p;
    GS_TRY(slot(c, SLOT_MLP_ACT, need, &p));
    float *f = static_cast<float *>(p);
    a->split = f;
    f += nsplit;
    a->X = f;
    f += static_cast<size_t>(rows) * MLP_F;
    a->Z1 = f; f += H;
    a->H1 = f; f += H;
    a->Z2 = f; f += H;
    a->H2 = f; f += H;
    a->Z3 = f; f += H;
    a->H3 = f; f += H;
    a->dH3 = f; f += H;
    a->dH2 = f; f += H;
    a->dH1 = f; f += H;
    a->Y = f; f += static_cast<size_t>(rows) * MLP_O;
    a->dY = f; f += static_cast<size_t>(rows) * MLP_O;
    a->loss = f;
    return make_success_error_status();
}

// forward of `rows` rows already in a.X (normalised features).  With T
// (targets, rows x MLP_O), the last layer's epilogue forms the MSE gradient
// dY = 2 (Y - T) / rows and the loss partials (a.loss) instead of Y.
static hipError_t forward(Context *c, const float *w, Acts &a, int rows, const float *T = nullptr)
{
    const float *W1 = w + MLP_OFF_W1, *b1 = w + MLP_OFF_B1, *W2 = w + MLP_OFF_W2,
                *b2 = w + MLP_OFF_B2, *W3 = w + MLP_OFF_W3, *b3 = w + MLP_OFF_B3,
                *W4 = w + MLP_OFF_W4, *b4 = w + MLP_OFF_B4;
    hipError_t e;
    GemmEpi ep{};
    ep.mode = 1;
    ep.bias = b1;
    ep.Z = a.Z1;
    if ((e = gemm(c, a.X, MLP_F, 1, W1, MLP_H, 1, rows, MLP_H, MLP_F, a.H1, MLP_H, ep))) return e;
    ep.bias = b2;
    ep.Z = a.Z2;
    if ((e = gemm(c, a.H1, MLP_H, 1, W2, MLP_H, 1, rows, MLP_H, MLP_H, a.H2, MLP_H, ep))) return e;
    ep.bias = b3;
    ep.Z = a.Z3;
    if ((e = gemm(c, a.H2, MLP_H, 1, W3, MLP_H, 1, rows, MLP_H, MLP_H, a.H3, MLP_H, ep))) return e;
    ep.bias = b4;
    ep.Z = nullptr;
    if (T) {
        ep.mode = 4;
        ep.T = T;
        ep.loss = a.loss;
        ep.inv_rows = 1.0f / static_cast<float>(rows);
        return gemm(c, a.H3, MLP_H, 1, W4, MLP_O, 1, rows, MLP_O, MLP_H, a.dY, MLP_O, ep);
    }
    ep.mode = 2;
    return gemm(c, a.H3, MLP_H, 1, W4, MLP_O, 1, rows, MLP_O, MLP_H, a.Y, MLP_O, ep);
}

// Gradient slices of sum over rows of |Y - T|^2 / rows into a.split (after
// forward(..., T)); returns the number of slices in *splits.
static hipError_t backward(Context *c, const float *w, Acts &a, int rows, int *splits)
{
    const float *W2 = w + MLP_OFF_W2, *W3 = w + MLP_OFF_W3, *W4 = w + MLP_OFF_W4;
    int S, kc;
    wgrad_slices(rows, &S, &kc);
    *splits = S;
    hipError_t e;
    GemmEpi ep{};
    ep.mode = 3;
    // layer 4: dW4 = H3^T dY, db4 = colsum dY; dZ3 = (dY W4^T) GELU'(Z3), into dH3
    if ((e = gemm_wgrad(c, a.H3, MLP_H, a.dY, MLP_O, rows, a.split, MLP_OFF_W4, MLP_OFF_B4, S, kc)))
        return e;
    ep.Z = a.Z3;
    if ((e = gemm(c, a.dY, MLP_O, 1, W4, 1, MLP_O, rows, MLP_H, MLP_O, a.dH3, MLP_H, ep))) return e;
    // layer 3
    if ((e = gemm_wgrad(c, a.H2, MLP_H, a.dH3, MLP_H, rows, a.split, MLP_OFF_W3, MLP_OFF_B3, S, kc)))
        return e;
    ep.Z = a.Z2;
    if ((e = gemm(c, a.dH3, MLP_H, 1, W3, 1, MLP_H, rows, MLP_H, MLP_H, a.dH2, MLP_H, ep))) return e;
    // layer 2
    if ((e = gemm_wgrad(c, a.H1, MLP_H, a.dH2, MLP_H, rows, a.split, MLP_OFF_W2, MLP_OFF_B2, S, kc)))
        return e;
    ep.Z = a.Z1;
    if ((e = gemm(c, a.dH2, MLP_H, 1, W2, 1, MLP_H, rows, MLP_H, MLP_H, a.dH1, MLP_H, ep))) return e;
    // layer 1
    return gemm_wgrad(c, a.X, MLP_F, a.dH1, MLP_H, rows, a.split, MLP_OFF_W1, MLP_OFF_B1, S, kc);
}

// ------------------------------------------------------------------ features
// Rotation-invariant inputs of one particle, from its coarse (monopole)
// acceleration a and the walk accumulators w (k_walk<false, true>, without
// the factor G), in the frame (e1, e2, e3) of mlp_local_frame:
//   0  G q.e1 / |a|      1  G q.e2 / |a|
//   2  G v.e1 / |a|      3  G v.e2 / |a|      4  G v.e3 / |a|
//   5  G s2 / |a|        6  G s3 / |a|        7  log10 |a|
//   8  log(1 + accepted nodes)                9  log(1 + particle interactions)
//   10 log10 max(G s3 / |a|, 1e-12)          11 log10 max(G s2 / |a|, 1e-12)
//   12 log10 max(G |q| / |a|, 1e-12)         13 |v| / s3 (0 if s3 = 0)
//   14, 15  0
// q = sum of the accepted nodes' quadrupole accelerations, v = sum
// (M l^3 / d^6) R, s2 = sum M l^2 / d^4, s3 = sum M l^3 / d^5.
__device__ inline void walk_vectors(const float *w, double q[3], double v[3])
{
    q[0] = w[0];
    q[1] = w[1];
    q[2] = w[2];
    v[0] = w[3];
    v[1] = w[4];
    v[2] = w[5];
}

__global__ void k_features(const double *__restrict__ a, const float *__restrict__ wf, int n,
                           double G, const float *__restrict__ mean, const float *__restrict__ stdv,
                           float *__restrict__ F)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double ai[3] = {a[3 * i], a[3 * i + 1], a[3 * i + 2]};
    const double an = sqrt(ai[0] * ai[0] + ai[1] * ai[1] + ai[2] * ai[2]);
    const float *w = wf + MLP_WALK_FEAT * static_cast<size_t>(i);
    double q[3], v[3], e1[3], e2[3], e3[3];
    walk_vectors(w, q, v);
    mlp_local_frame(ai, q, v, e1, e2, e3);
    const double g = an > 0.0 ? G / an : 0.0;
    auto dot = [](const double *x, const double *y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
    const double s2 = g * w[6], s3 = g * w[7];
    const double qn = g * sqrt(dot(q, q));
    float f[MLP_F];
    f[0] = static_cast<float>(g * dot(q, e1));
    f[1] = static_cast<float>(g * dot(q, e2));
    f[2] = static_cast<float>(g * dot(v, e1));
    f[3] = static_cast<float>(g * dot(v, e2));
    f[4] = static_cast<float>(g * dot(v, e3));
    f[5] = static_cast<float>(s2);
    f[6] = static_cast<float>(s3);
    f[7] = an > 0.0 ? static_cast<float>(log10(an)) : -30.0f;
    f[8] = static_cast<float>(log1p(static_cast<double>(w[8])));
    f[9] = static_cast<float>(log1p(static_cast<double>(w[9])));
    f[10] = static_cast<float>(log10(fmax(s3, 1e-12)));
    f[11] = static_cast<float>(log10(fmax(s2, 1e-12)));
    f[12] = static_cast<float>(log10(fmax(qn, 1e-12)));
    // the walk's octupole sum in the local frame, in units of the target's
    // scale s3 (both without G): the physics of the next order, whose
    // residual the network predicts
    const double oc[3] = {w[10], w[11], w[12]};
    const double is3 = w[7] > 0.0f ? 1.0 / static_cast<double>(w[7]) : 0.0;
    f[13] = static_cast<float>(dot(oc, e1) * is3);
    f[14] = static_cast<float>(dot(oc, e2) * is3);
    f[15] = static_cast<float>(dot(oc, e3) * is3);
    float *o = F + MLP_F * static_cast<size_t>(i);
#pragma unroll
    for (int k = 0; k < MLP_F; k++) o[k] = mean ? (f[k] - mean[k]) / stdv[k] : f[k];
}

// ------------------------------------------------------------------ fused inference
// One kernel for the whole correction of method BARNES_HUT_LEARNED: each
// wavefront takes 16 particles at a time through the four layers with the
// bf16 weights resident in LDS (transposed to (outputs, inputs) so that a
// lane's eight k-values of a B fragment are one 16-byte read), keeps the
// 16 x 128 activations of its particles in its own LDS tile between layers,
// and finally applies
//   a += G (q + o) + G s3 * tscale * (y0 e1 + y1 e2 + y2 e3)
// in the particle's local frame.  Rows of the LDS tiles are padded by 8
// bf16 (16 B) so that the 16 rows a fragment read touches fall in
// different banks.
constexpr int MI_WAVES = 8;
constexpr int MI_KP1 = 32 + 8;        // layer-1 weight rows: K = 16 padded to 32
constexpr int MI_KPH = MLP_H + 8;     // 128-wide rows

struct MiLds {
    __bf16 w1[MLP_H * MI_KP1];
    __bf16 w2[MLP_H * MI_KPH];
    __bf16 w3[MLP_H * MI_KPH];
    __bf16 w4[16 * MI_KPH];
    float b1[MLP_H], b2[MLP_H], b3[MLP_H], b4[16];
    __bf16 h[MI_WAVES][16 * MI_KPH];
    float y[MI_WAVES][16 * 4];
};

// Hidden layer: 16 rows of h (K = 128) times W (LDS, [n][k]) plus bias,
// GELU, written back into h.
__device__ __forceinline__ void mi_hidden(__bf16 *h, const __bf16 *w, const float *b, int r, int q)
{
    bf16x8 a[4];
#pragma unroll
    for (int ks = 0; ks < 4; ks++) a[ks] = lds8(h + r * MI_KPH + ks * 32 + 8 * q);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
#pragma unroll
    for (int nt = 0; nt < MLP_H / 16; nt++) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        const __bf16 *wn = w + (nt * 16 + r) * MI_KPH + 8 * q;
#pragma unroll
        for (int ks = 0; ks < 4; ks++)
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ks], lds8(wn + ks * 32), acc, 0, 0, 0);
        const int col = nt * 16 + r;
        const float bc = b[col];
#pragma unroll
        for (int j = 0; j < 4; j++) h[(4 * q + j) * MI_KPH + col] = to_bf16(gelu(acc[j] + bc));
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
}

__global__ __launch_bounds__(64 * MI_WAVES) void k_mlp_infer(const float *__restrict__ F,
                                                              const float *__restrict__ wf,
                                                              const float *__restrict__ w, int n,
                                                              double G, float tscale,
                                                              double *__restrict__ a)
{
    extern __shared__ __align__(16) unsigned char mi_smem[];
    MiLds &S = *reinterpret_cast<MiLds *>(mi_smem);
    const int tid = threadIdx.x;
    // weights: FP32 master (row-major (k, n)) -> bf16 transposed in LDS
    for (int i = tid; i < MLP_H * MI_KP1; i += 64 * MI_WAVES) {
        const int nn = i / MI_KP1, k = i % MI_KP1;
        S.w1[i] = to_bf16(k < MLP_F ? w[MLP_OFF_W1 + k * MLP_H + nn] : 0.0f);
    }
    for (int i = tid; i < MLP_H * MLP_H; i += 64 * MI_WAVES) {
        const int k = i / MLP_H, nn = i % MLP_H;
        S.w2[nn * MI_KPH + k] = to_bf16(w[MLP_OFF_W2 + i]);
        S.w3[nn * MI_KPH + k] = to_bf16(w[MLP_OFF_W3 + i]);
    }
    for (int i = tid; i < 16 * MI_KPH; i += 64 * MI_WAVES) {
        const int nn = i / MI_KPH, k = i % MI_KPH;
        S.w4[i] = to_bf16(nn < MLP_O && k < MLP_H ? w[MLP_OFF_W4 + k * MLP_O + nn] : 0.0f);
    }
    for (int i = tid; i < MLP_H; i += 64 * MI_WAVES) {
        S.b1[i] = w[MLP_OFF_B1 + i];
        S.b2[i] = w[MLP_OFF_B2 + i];
        S.b3[i] = w[MLP_OFF_B3 + i];
    }
    if (tid < 16) S.b4[tid] = tid < MLP_O ? w[MLP_OFF_B4 + tid] : 0.0f;
    __syncthreads();

    const int lane = tid & 63, wave = tid >> 6;
    const int r = lane & 15, q = lane >> 4;
    __bf16 *h = S.h[wave];
    float *y = S.y[wave];
    const int ntiles = (n + 15) / 16;
    for (int tile = blockIdx.x * MI_WAVES + wave; tile < ntiles; tile += gridDim.x * MI_WAVES) {
        const int row0 = tile * 16;
        // layer 1: K = 16 features (lanes q = 2, 3 carry the zero padding)
        bf16x8 x;
        {
            const int row = row0 + r;
            float4 u0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), u1 = u0;
            if (q < 2 && row < n) {
                const float4 *src = reinterpret_cast<const float4 *>(F + static_cast<size_t>(row) * MLP_F + 8 * q);
                u0 = src[0];
                u1 = src[1];
            }
            x[0] = to_bf16(u0.x); x[1] = to_bf16(u0.y); x[2] = to_bf16(u0.z); x[3] = to_bf16(u0.w);
            x[4] = to_bf16(u1.x); x[5] = to_bf16(u1.y); x[6] = to_bf16(u1.z); x[7] = to_bf16(u1.w);
        }
#pragma unroll
        for (int nt = 0; nt < MLP_H / 16; nt++) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                x, lds8(S.w1 + (nt * 16 + r) * MI_KP1 + 8 * q), acc, 0, 0, 0);
            const int col = nt * 16 + r;
            const float bc = S.b1[col];
#pragma unroll
            for (int j = 0; j < 4; j++) h[(4 * q + j) * MI_KPH + col] = to_bf16(gelu(acc[j] + bc));
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        mi_hidden(h, S.w2, S.b2, r, q);
        mi_hidden(h, S.w3, S.b3, r, q);
        // layer 4: 3 outputs (columns 3..15 of the tile are zero weights)
        {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ks = 0; ks < 4; ks++)
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    lds8(h + r * MI_KPH + ks * 32 + 8 * q), lds8(S.w4 + r * MI_KPH + ks * 32 + 8 * q),
                    acc, 0, 0, 0);
            if (r < MLP_O) {
#pragma unroll
                for (int j = 0; j < 4; j++) y[(4 * q + j) * 4 + r] = acc[j] + S.b4[r];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        // correction of the tile's 16 particles, one lane each
        if (lane < 16 && row0 + lane < n) {
            const size_t i = static_cast<size_t>(row0 + lane);
            const double ai[3] = {a[3 * i], a[3 * i + 1], a[3 * i + 2]};
            const float *wi = wf + MLP_WALK_FEAT * i;
            double qv[3], vv[3], e1[3], e2[3], e3[3];
            walk_vectors(wi, qv, vv);
            mlp_local_frame(ai, qv, vv, e1, e2, e3);
            const double s = G * static_cast<double>(wi[7]) * static_cast<double>(tscale);
            const double y0 = y[lane * 4], y1 = y[lane * 4 + 1], y2 = y[lane * 4 + 2];
            const double ov[3] = {wi[10], wi[11], wi[12]};
            for (int d = 0; d < 3; d++)
                a[3 * i + d] = ai[d] + G * (qv[d] + ov[d]) + s * (y0 * e1[d] + y1 * e2[d] + y2 * e3[d]);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    }
}

// ------------------------------------------------------------------ host API pieces
ErrorStatus mlp_forward_rows(Context *c, const float *X, int rows, float *Y)
{
    GS_TRY(mlp_ensure(c));
    Acts a;
    GS_TRY(acts(c, rows, &a));
    GS_HIP(hipMemcpyAsync(a.X, X, sizeof(float) * rows * MLP_F, hipMemcpyDeviceToDevice, c->stream));
    GS_HIP(forward(c, g_model.w, a, rows));
    GS_HIP(hipMemcpyAsync(Y, a.Y, sizeof(float) * rows * MLP_O, hipMemcpyDeviceToDevice, c->stream));
    return make_success_error_status();
}

ErrorStatus mlp_gradients(Context *c, const float *X, const float *T, int rows, float *grad,
                          double *loss)
{
    GS_TRY(mlp_ensure(c));
    Acts a;
    GS_TRY(acts(c, rows, &a));
    int splits = 0, np;
    if (fused_rows(rows)) {
        char *img;
        GS_TRY(weight_image(c, g_model.w, &img));
        GS_TRY(train_slices(c, X, T, nullptr, rows, img, a.split, a.loss, &splits));
        np = splits * TS_WAVES;
    } else {
        GS_HIP(hipMemcpyAsync(a.X, X, sizeof(float) * rows * MLP_F, hipMemcpyDeviceToDevice, c->stream));
        GS_HIP(forward(c, g_model.w, a, rows, T));
        GS_HIP(backward(c, g_model.w, a, rows, &splits));
        np = gemm_loss_parts(rows, MLP_O);
    }
    hipLaunchKernelGGL(k_reduce_grads, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, c->stream,
                       a.split, splits, grad, 0, nullptr, nullptr, nullptr, 0.0f, 0.0f, 0.0f, 0.0f,
                       0.0f, 0.0f, nullptr);
    GS_HIP(hipGetLastError());
    std::vector<float> part(np);
    GS_HIP(hipMemcpyAsync(part.data(), a.loss, sizeof(float) * np, hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    double l = 0.0;
    for (float p : part) l += p;
    *loss = l / rows;
    return make_success_error_status();
}

ErrorStatus mlp_train(Context *c, const float *X, const float *T, int S, int epochs, int batch,
                      double lr, unsigned seed, double *final_loss)
{
    GS_TRY(mlp_ensure(c));
    MlpModel &Mo = g_model;
    if (batch <= 0 || batch > S) batch = S;
    // He-style initialisation from a deterministic LCG (host), zero biases
    {
        std::vector<float> w(MLP_NPARAM, 0.0f);
        unsigned long long st = 0x9E3779B97F4A7C15ull ^ seed;
        auto uni = [&]() {
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            return static_cast<double>(st >> 11) * (1.0 / 9007199254740992.0);
        };
        auto fill = [&](int off, int fan_in, int count) {
            const double sc = sqrt(6.0 / fan_in);
            for (int k = 0; k < count; k++) w[off + k] = static_cast<float>((2.0 * uni() - 1.0) * sc);
        };
        fill(MLP_OFF_W1, MLP_F, MLP_F * MLP_H);
        fill(MLP_OFF_W2, MLP_H, MLP_H * MLP_H);
        fill(MLP_OFF_W3, MLP_H, MLP_H * MLP_H);
        fill(MLP_OFF_W4, MLP_H, MLP_H * MLP_O);
        for (int k = 0; k < MLP_H * MLP_O; k++) w[MLP_OFF_W4 + k] *= 0.1f;
        GS_HIP(hipMemcpyAsync(Mo.w, w.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice,
                              c->stream));
        GS_HIP(hipMemsetAsync(Mo.m, 0, sizeof(float) * MLP_NPARAM, c->stream));
        GS_HIP(hipMemsetAsync(Mo.v, 0, sizeof(float) * MLP_NPARAM, c->stream));
    }
    Acts a;
    GS_TRY(acts(c, batch, &a));
    const int steps_per_epoch = S / batch;
    const long long total = static_cast<long long>(epochs) * steps_per_epoch;
    if (static_cast<long long>(epochs) * S > (1LL << 31) - 1)
        return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR, "too many training steps");
    // device: the batch targets and every epoch's permutation of the rows
    const size_t tb_bytes = (sizeof(float) * batch * MLP_O + 255) / 256 * 256;
    void *p;
    GS_TRY(slot(c, SLOT_MLP_GRAD, tb_bytes + sizeof(int) * static_cast<size_t>(epochs) * S, &p));
    float *Tb = static_cast<float *>(p);
    int *perm = reinterpret_cast<int *>(static_cast<char *>(p) + tb_bytes);
    {
        std::vector<int> hperm(static_cast<size_t>(epochs) * S);
        std::vector<int> cur(S);
        for (int i = 0; i < S; i++) cur[i] = i;
        unsigned long long st = 0x2545F4914F6CDD1Dull ^ (seed * 7919ull);
        for (int ep = 0; ep < epochs; ep++) {
            for (int i = S - 1; i > 0; i--) {  // Fisher-Yates, continued from the last epoch's order
                st = st * 6364136223846793005ull + 1442695040888963407ull;
                const int j = static_cast<int>((st >> 33) % static_cast<unsigned long long>(i + 1));
                std::swap(cur[i], cur[j]);
            }
            std::copy(cur.begin(), cur.end(), hperm.begin() + static_cast<size_t>(ep) * S);
        }
        GS_HIP(hipMemcpyAsync(perm, hperm.data(), sizeof(int) * hperm.size(), hipMemcpyHostToDevice,
                              c->stream));
        GS_HIP(hipStreamSynchronize(c->stream));  // the host vector goes out of scope
    }
    const float b1 = 0.9f, b2 = 0.999f;
    const bool fused = fused_rows(batch);
    char *img = nullptr;
    if (fused) GS_TRY(weight_image(c, Mo.w, &img));
    for (long long t = 0; t < total; t++) {
        const int *idx = perm + (t / steps_per_epoch) * S + (t % steps_per_epoch) * batch;
        int splits = 0;
        if (fused) {
            // one kernel: batch gather, forward, MSE gradient, backward
            GS_TRY(train_slices(c, X, T, idx, batch, img, a.split, a.loss, &splits));
        } else {
            hipLaunchKernelGGL(k_gather_batch, dim3((batch * (MLP_F + MLP_O) + 255) / 256), dim3(256),
                               0, c->stream, X, T, idx, batch, a.X, Tb);
            GS_HIP(forward(c, Mo.w, a, batch, Tb));
            GS_HIP(backward(c, Mo.w, a, batch, &splits));
        }
        // cosine decay to 5% of lr
        const double frac = static_cast<double>(t + 1) / static_cast<double>(total);
        const float lr_t = static_cast<float>(lr * (0.05 + 0.95 * 0.5 * (1.0 + cos(M_PI * frac))));
        const float bc1 = 1.0f - powf(b1, static_cast<float>(t + 1));
        const float bc2 = 1.0f - powf(b2, static_cast<float>(t + 1));
        hipLaunchKernelGGL(k_reduce_grads, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, c->stream,
                           a.split, splits, Mo.g, 1, Mo.w, Mo.m, Mo.v, lr_t, b1, b2, bc1, bc2, 1e-8f, img);
        GS_HIP(hipGetLastError());
    }
    // the last step's loss
    double last = 0.0;
    if (total > 0) {
        const int np = fused ? (batch + TS_ROWS - 1) / TS_ROWS * TS_WAVES : gemm_loss_parts(batch, MLP_O);
        std::vector<float> part(np);
        GS_HIP(hipMemcpyAsync(part.data(), a.loss, sizeof(float) * np, hipMemcpyDeviceToHost,
                              c->stream));
        GS_HIP(hipStreamSynchronize(c->stream));
        for (float v : part) last += v;
        last /= batch;
    }
    Mo.trained = true;
    if (final_loss) *final_loss = last;
    return make_success_error_status();
}

ErrorStatus mlp_features(Context *c, const double *a, const float *walkfeat, int n, double G,
                         bool normalise, float *F)
{
    GS_TRY(mlp_ensure(c));
    float *mean = nullptr, *stdv = nullptr;
    if (normalise) {
        void *p;
        GS_TRY(slot(c, SLOT_MLP_W, sizeof(float) * 2 * MLP_F, &p));
        mean = static_cast<float *>(p);
        stdv = mean + MLP_F;
        GS_HIP(hipMemcpyAsync(mean, g_model.fmean, sizeof(float) * MLP_F, hipMemcpyHostToDevice, c->stream));
        GS_HIP(hipMemcpyAsync(stdv, g_model.fstd, sizeof(float) * MLP_F, hipMemcpyHostToDevice, c->stream));
    }
    hipLaunchKernelGGL(k_features, dim3((n + 255) / 256), dim3(256), 0, c->stream, a, walkfeat, n,
                       G, mean, stdv, F);
    GS_HIP(hipGetLastError());
    if (normalise) GS_HIP(hipStreamSynchronize(c->stream));  // the host stats buffer is reused
    return make_success_error_status();
}

// Corrects a (n particles) in place from normalised features F and the walk
// accumulators wf.
ErrorStatus mlp_apply(Context *c, const float *F, const float *wf, double G, int n, double *a)
{
    GS_TRY(mlp_ensure(c));
    static bool attr_set = false;
    const size_t lds = sizeof(MiLds);
    if (!attr_set) {
        GS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_mlp_infer),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
        attr_set = true;
    }
    const int ntiles = (n + 15) / 16;
    int grid = (ntiles + MI_WAVES - 1) / MI_WAVES;
    if (grid > c->cu_count) grid = c->cu_count;  // one resident workgroup per CU, looping
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(64 * MI_WAVES), lds, c->stream, F, wf,
                       g_model.w, n, G, g_model.tscale, a);
    GS_HIP(hipGetLastError());
    return make_success_error_status();
}

}  // namespace gs
