// mlp.hip -- the learned Barnes-Hut correction: a 16 -> 128 -> 128 -> 128 -> 3
// MLP with tanh-GELU, trained and evaluated on the GPU with BF16 MFMA
// (v_mfma_f32_16x16x32_bf16) and FP32 accumulation.
//
//   k_gemm        C = op(A) op(B) (+ epilogues: bias, bias+GELU with the
//                 pre-activation kept for the backward pass).  One wavefront
//                 per 16x16 output tile, K in steps of 32; operands are read
//                 as FP32 and rounded to BF16 (nearest even) into the MFMA
//                 fragments.  Fragment layout (16x16x32): lane l holds
//                 A[l%16][k0 + 8(l/16) + 0..7], B[k0 + 8(l/16) + 0..7][l%16];
//                 the result D[4(l/16) + 0..3][l%16].
//   k_gelu_bwd    dZ = dH * GELU'(Z)
//   k_colsum      bias gradients (column sums, fixed order)
//   k_mse         dY = 2 (Y - T) / B and per-block loss partials
//   k_adam        Adam on the FP32 master weights
//   k_features    per-particle input features from the coarse walk
//   k_mlp_infer   fused inference + correction for method BARNES_HUT_LEARNED:
//                 a = a_coarse + G (q + o) + G s3 * tscale * (y in the local frame)
//                 (q, o: the walk's quadrupole and octupole sums, the physics
//                 part of the predicted residual; s3 = sum M l^3 / d^5, the
//                 scale of the next multipole orders)
#include <cmath>
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

__device__ __forceinline__ float gelu(float z)
{
    const float c = 0.7978845608028654f;  // sqrt(2 / pi)
    const float u = c * (z + 0.044715f * z * z * z);
    return 0.5f * z * (1.0f + tanhf(u));
}

__device__ __forceinline__ float gelu_grad(float z)
{
    const float c = 0.7978845608028654f;
    const float u = c * (z + 0.044715f * z * z * z);
    const float t = tanhf(u);
    return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * c * (1.0f + 0.134145f * z * z);
}

// C(m, n) = sum_k A(m, k) B(k, n); A(m, k) = A[m * sam + k * sak], B likewise.
// mode 0: C = acc; 1: Z = acc + bias, C = GELU(Z); 2: C = acc + bias.
__global__ __launch_bounds__(256) void k_gemm(const float *__restrict__ A, long long sam,
                                              long long sak, const float *__restrict__ B,
                                              long long sbk, long long sbn, int M, int N, int K,
                                              float *__restrict__ C, long long ldc,
                                              const float *__restrict__ bias,
                                              float *__restrict__ Z, int mode)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int tm = (blockIdx.x * 2 + (wave >> 1)) * 16;
    const int tn = (blockIdx.y * 2 + (wave & 1)) * 16;
    if (tm >= M || tn >= N) return;
    const int r = lane & 15, q = lane >> 4;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const int am = tm + r, bn = tn + r;
    for (int k0 = 0; k0 < K; k0 += 32) {
        bf16x8 a, b;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int k = k0 + 8 * q + j;
            const float av = (am < M && k < K) ? A[am * sam + k * sak] : 0.0f;
            const float bv = (bn < N && k < K) ? B[k * sbk + bn * sbn] : 0.0f;
            a[j] = to_bf16(av);
            b[j] = to_bf16(bv);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    const int cn = tn + r;
    if (cn >= N) return;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int cm = tm + 4 * q + j;
        if (cm >= M) continue;
        float v = acc[j];
        if (mode >= 1) v += bias[cn];
        if (mode == 1) {
            Z[cm * ldc + cn] = v;
            v = gelu(v);
        }
        C[cm * ldc + cn] = v;
    }
}

static hipError_t gemm(Context *c, const float *A, long long sam, long long sak, const float *B,
                       long long sbk, long long sbn, int M, int N, int K, float *C, long long ldc,
                       const float *bias, float *Z, int mode)
{
    dim3 grid((M + 31) / 32, (N + 31) / 32);
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, c->stream, A, sam, sak, B, sbk, sbn, M, N, K,
                       C, ldc, bias, Z, mode);
    return hipGetLastError();
}

__global__ void k_gelu_bwd(float *__restrict__ dH, const float *__restrict__ Z, long long n)
{
    const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
    if (i < n) dH[i] *= gelu_grad(Z[i]);
}

// out[n] = sum over rows of X[row * ld + n], rows in fixed order per block.
__global__ __launch_bounds__(256) void k_colsum(const float *__restrict__ X, int rows, int ld,
                                                float *__restrict__ out)
{
    __shared__ float red[256];
    const int n = blockIdx.x;
    float s = 0.0f;
    for (int r = threadIdx.x; r < rows; r += 256) s += X[static_cast<long long>(r) * ld + n];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[n] = red[0];
}

// dY = 2 (Y - T) / B ; loss partials per block
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ Y, const float *__restrict__ T,
                                             int n, float inv_b, float *__restrict__ dY,
                                             float *__restrict__ part)
{
    __shared__ float red[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    float l = 0.0f;
    if (i < n) {
        const float d = Y[i] - T[i];
        dY[i] = 2.0f * d * inv_b;
        l = d * d;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int n, float lr, float b1, float b2, float bc1,
                       float bc2, float eps)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
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
};

static ErrorStatus acts(Context *c, int rows, Acts *a)
{
    const size_t H = static_cast<size_t>(rows) * MLP_H;
    const size_t need = sizeof(float) * (static_cast<size_t>(rows) * MLP_F + 9 * H +
                                         2 * static_cast<size_t>(rows) * MLP_O + rows + 1024);
    void *p;
    GS_TRY(slot(c, SLOT_MLP_ACT, need, &p));
    float *f = static_cast<float *>(p);
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

// forward of `rows` rows already in a.X (normalised features)
static hipError_t forward(Context *c, const float *w, Acts &a, int rows)
{
    const float *W1 = w + MLP_OFF_W1, *b1 = w + MLP_OFF_B1, *W2 = w + MLP_OFF_W2,
                *b2 = w + MLP_OFF_B2, *W3 = w + MLP_OFF_W3, *b3 = w + MLP_OFF_B3,
                *W4 = w + MLP_OFF_W4, *b4 = w + MLP_OFF_B4;
    hipError_t e;
    if ((e = gemm(c, a.X, MLP_F, 1, W1, MLP_H, 1, rows, MLP_H, MLP_F, a.H1, MLP_H, b1, a.Z1, 1)))
        return e;
    if ((e = gemm(c, a.H1, MLP_H, 1, W2, MLP_H, 1, rows, MLP_H, MLP_H, a.H2, MLP_H, b2, a.Z2, 1)))
        return e;
    if ((e = gemm(c, a.H2, MLP_H, 1, W3, MLP_H, 1, rows, MLP_H, MLP_H, a.H3, MLP_H, b3, a.Z3, 1)))
        return e;
    return gemm(c, a.H3, MLP_H, 1, W4, MLP_O, 1, rows, MLP_O, MLP_H, a.Y, MLP_O, b4, nullptr, 2);
}

// gradients of sum over rows of |Y - T|^2 / rows into g (requires forward)
static hipError_t backward(Context *c, const float *w, float *g, Acts &a, const float *T, int rows,
                           int *nparts)
{
    const float *W2 = w + MLP_OFF_W2, *W3 = w + MLP_OFF_W3, *W4 = w + MLP_OFF_W4;
    const int nY = rows * MLP_O;
    *nparts = (nY + 255) / 256;
    hipLaunchKernelGGL(k_mse, dim3(*nparts), dim3(256), 0, c->stream, a.Y, T, nY,
                       1.0f / static_cast<float>(rows), a.dY, a.loss);
    hipError_t e;
    // layer 4: dW4 = H3^T dY, db4 = colsum dY, dH3 = dY W4^T
    if ((e = gemm(c, a.H3, 1, MLP_H, a.dY, MLP_O, 1, MLP_H, MLP_O, rows, g + MLP_OFF_W4, MLP_O,
                  nullptr, nullptr, 0)))
        return e;
    hipLaunchKernelGGL(k_colsum, dim3(MLP_O), dim3(256), 0, c->stream, a.dY, rows, MLP_O,
                       g + MLP_OFF_B4);
    if ((e = gemm(c, a.dY, MLP_O, 1, W4, 1, MLP_O, rows, MLP_H, MLP_O, a.dH3, MLP_H, nullptr,
                  nullptr, 0)))
        return e;
    const long long nH = static_cast<long long>(rows) * MLP_H;
    const dim3 gH((nH + 255) / 256);
    // layer 3
    hipLaunchKernelGGL(k_gelu_bwd, gH, dim3(256), 0, c->stream, a.dH3, a.Z3, nH);
    if ((e = gemm(c, a.H2, 1, MLP_H, a.dH3, MLP_H, 1, MLP_H, MLP_H, rows, g + MLP_OFF_W3, MLP_H,
                  nullptr, nullptr, 0)))
        return e;
    hipLaunchKernelGGL(k_colsum, dim3(MLP_H), dim3(256), 0, c->stream, a.dH3, rows, MLP_H,
                       g + MLP_OFF_B3);
    if ((e = gemm(c, a.dH3, MLP_H, 1, W3, 1, MLP_H, rows, MLP_H, MLP_H, a.dH2, MLP_H, nullptr,
                  nullptr, 0)))
        return e;
    // layer 2
    hipLaunchKernelGGL(k_gelu_bwd, gH, dim3(256), 0, c->stream, a.dH2, a.Z2, nH);
    if ((e = gemm(c, a.H1, 1, MLP_H, a.dH2, MLP_H, 1, MLP_H, MLP_H, rows, g + MLP_OFF_W2, MLP_H,
                  nullptr, nullptr, 0)))
        return e;
    hipLaunchKernelGGL(k_colsum, dim3(MLP_H), dim3(256), 0, c->stream, a.dH2, rows, MLP_H,
                       g + MLP_OFF_B2);
    if ((e = gemm(c, a.dH2, MLP_H, 1, W2, 1, MLP_H, rows, MLP_H, MLP_H, a.dH1, MLP_H, nullptr,
                  nullptr, 0)))
        return e;
    // layer 1
    hipLaunchKernelGGL(k_gelu_bwd, gH, dim3(256), 0, c->stream, a.dH1, a.Z1, nH);
    if ((e = gemm(c, a.X, 1, MLP_F, a.dH1, MLP_H, 1, MLP_F, MLP_H, rows, g + MLP_OFF_W1, MLP_H,
                  nullptr, nullptr, 0)))
        return e;
    hipLaunchKernelGGL(k_colsum, dim3(MLP_H), dim3(256), 0, c->stream, a.dH1, rows, MLP_H,
                       g + MLP_OFF_B1);
    return hipGetLastError();
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

// GELU (tanh form) as z * sigmoid(2u): 0.5 (1 + tanh(u)) = 1 / (1 + exp(-2u))
__device__ __forceinline__ float gelu_fast(float z)
{
    const float u = 0.7978845608028654f * (z + 0.044715f * z * z * z);
    return z / (1.0f + __expf(-2.0f * u));
}

__device__ __forceinline__ bf16x8 lds8(const __bf16 *p)
{
    return *reinterpret_cast<const bf16x8 *>(p);
}

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
        for (int j = 0; j < 4; j++) h[(4 * q + j) * MI_KPH + col] = to_bf16(gelu_fast(acc[j] + bc));
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
            for (int j = 0; j < 4; j++) h[(4 * q + j) * MI_KPH + col] = to_bf16(gelu_fast(acc[j] + bc));
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
    GS_HIP(hipMemcpyAsync(a.X, X, sizeof(float) * rows * MLP_F, hipMemcpyDeviceToDevice, c->stream));
    GS_HIP(forward(c, g_model.w, a, rows));
    int np = 0;
    GS_HIP(backward(c, g_model.w, grad, a, T, rows, &np));
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
    void *p;
    GS_TRY(slot(c, SLOT_MLP_GRAD, sizeof(float) * batch * MLP_O + sizeof(int) * S + 256, &p));
    float *Tb = static_cast<float *>(p);
    int *perm = reinterpret_cast<int *>(Tb + batch * MLP_O);
    std::vector<int> hperm(S);
    for (int i = 0; i < S; i++) hperm[i] = i;
    unsigned long long st = 0x2545F4914F6CDD1Dull ^ (seed * 7919ull);
    const int steps_per_epoch = S / batch;
    int step = 0;
    double last = 0.0;
    const float b1 = 0.9f, b2 = 0.999f;
    for (int ep = 0; ep < epochs; ep++) {
        for (int i = S - 1; i > 0; i--) {  // Fisher-Yates
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            const int j = static_cast<int>((st >> 33) % static_cast<unsigned long long>(i + 1));
            std::swap(hperm[i], hperm[j]);
        }
        GS_HIP(hipMemcpyAsync(perm, hperm.data(), sizeof(int) * S, hipMemcpyHostToDevice, c->stream));
        double ep_loss = 0.0;
        for (int s = 0; s < steps_per_epoch; s++) {
            const int *idx = perm + s * batch;
            hipLaunchKernelGGL(k_gather_rows, dim3((batch * MLP_F + 255) / 256), dim3(256), 0,
                               c->stream, X, idx, batch, MLP_F, a.X);
            hipLaunchKernelGGL(k_gather_rows, dim3((batch * MLP_O + 255) / 256), dim3(256), 0,
                               c->stream, T, idx, batch, MLP_O, Tb);
            GS_HIP(forward(c, Mo.w, a, batch));
            int np = 0;
            GS_HIP(backward(c, Mo.w, Mo.g, a, Tb, batch, &np));
            step++;
            // cosine decay to 5% of lr
            const double frac = static_cast<double>(step) / (epochs * steps_per_epoch);
            const float lr_t = static_cast<float>(lr * (0.05 + 0.95 * 0.5 * (1.0 + cos(M_PI * frac))));
            const float bc1 = 1.0f - powf(b1, static_cast<float>(step));
            const float bc2 = 1.0f - powf(b2, static_cast<float>(step));
            hipLaunchKernelGGL(k_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, c->stream,
                               Mo.w, Mo.g, Mo.m, Mo.v, MLP_NPARAM, lr_t, b1, b2, bc1, bc2, 1e-8f);
            GS_HIP(hipGetLastError());
            if (s == steps_per_epoch - 1) {
                std::vector<float> part(np);
                GS_HIP(hipMemcpyAsync(part.data(), a.loss, sizeof(float) * np,
                                      hipMemcpyDeviceToHost, c->stream));
                GS_HIP(hipStreamSynchronize(c->stream));
                for (float v : part) ep_loss += v;
                last = ep_loss / batch;
            }
        }
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
