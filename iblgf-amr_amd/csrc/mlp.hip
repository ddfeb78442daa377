// gfx950 kernels of the learned preconditioner's MLP (training and inference).
// Sizes are small (hidden 128, ~10^4 points per batch), so one tiled FP32
// GEMM kernel with LDS tiles of 64x64 and a 4x4 register block per thread
// serves the forward pass, the backward data pass and the weight gradients
// (transposed operand access selected at compile time).  Reductions run in a
// fixed order, so training is deterministic.
#include "mlp.hpp"

#include <cmath>
#include <stdexcept>

namespace iblgf_amd {
namespace {

constexpr int kThreads = 256;

__device__ inline float bf16_to_f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }
__device__ inline uint16_t f_to_bf16(float f)
{
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);  // NaN stays NaN
    return static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

template<bool TA, bool TB, bool BF>
__global__ __launch_bounds__(kThreads) void k_gemm(int M, int N, int K, const float* __restrict__ A, int lda,
    const void* __restrict__ Bv, int ldb, const float* __restrict__ bias, int act, const float* __restrict__ P,
    float* __restrict__ C)
{
    __shared__ float As[16][64 + 4];
    __shared__ float Bs[16][64 + 4];
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    float     acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.f;
    for (int k0 = 0; k0 < K; k0 += 16)
    {
        for (int e = threadIdx.x; e < 1024; e += kThreads)
        {
            int mm, kk;
            if (TA)
            {
                mm = e % 64;
                kk = e / 64;
            }
            else
            {
                kk = e % 16;
                mm = e / 16;
            }
            const int gm = m0 + mm, gk = k0 + kk;
            float     v = 0.f;
            if (gm < M && gk < K) v = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
            As[kk][mm] = v;
        }
        for (int e = threadIdx.x; e < 1024; e += kThreads)
        {
            int kk, nn;
            if (TB)
            {
                kk = e % 16;
                nn = e / 16;
            }
            else
            {
                nn = e % 64;
                kk = e / 64;
            }
            const int gk = k0 + kk, gn = n0 + nn;
            float     v = 0.f;
            if (gk < K && gn < N)
            {
                const size_t idx = TB ? (size_t)gn * ldb + gk : (size_t)gk * ldb + gn;
                v = BF ? bf16_to_f(static_cast<const uint16_t*>(Bv)[idx]) : static_cast<const float*>(Bv)[idx];
            }
            Bs[kk][nn] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk)
        {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            const int m = m0 + ty * 4 + i, n = n0 + tx * 4 + j;
            if (m >= M || n >= N) continue;
            float v = acc[i][j];
            if (bias) v += bias[n];
            if (act == 1) v = tanhf(v);
            if (P)
            {
                const float p = P[(size_t)m * N + n];
                v *= (1.f - p * p);
            }
            C[(size_t)m * N + n] = v;
        }
}

__global__ __launch_bounds__(kThreads) void k_features(int npts, int K, const double* __restrict__ r,
    const int* __restrict__ nbr, const float* __restrict__ frame, const float* __restrict__ rel, float inv_s,
    float* __restrict__ X)
{
    const int t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= npts * K) return;
    const int    i = t / K, j = t - i * K;
    const int    q = nbr[(size_t)i * K + j];
    const float* R = frame + 9 * (size_t)i;
    const float  v0 = (float)r[3 * (size_t)q], v1 = (float)r[3 * (size_t)q + 1], v2 = (float)r[3 * (size_t)q + 2];
    float*       x = X + (size_t)i * 6 * K;
    x[3 * j + 0] = (R[0] * v0 + R[1] * v1 + R[2] * v2) * inv_s;
    x[3 * j + 1] = (R[3] * v0 + R[4] * v1 + R[5] * v2) * inv_s;
    x[3 * j + 2] = (R[6] * v0 + R[7] * v1 + R[8] * v2) * inv_s;
    const float* p = rel + ((size_t)i * K + j) * 3;
    x[3 * K + 3 * j + 0] = p[0];
    x[3 * K + 3 * j + 1] = p[1];
    x[3 * K + 3 * j + 2] = p[2];
}

__global__ __launch_bounds__(kThreads) void k_targets(int npts, const double* __restrict__ e,
    const float* __restrict__ frame, float inv, float* __restrict__ T)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= npts) return;
    const float* R = frame + 9 * (size_t)i;
    const float  v0 = (float)e[3 * (size_t)i], v1 = (float)e[3 * (size_t)i + 1], v2 = (float)e[3 * (size_t)i + 2];
    T[3 * (size_t)i + 0] = (R[0] * v0 + R[1] * v1 + R[2] * v2) * inv;
    T[3 * (size_t)i + 1] = (R[3] * v0 + R[4] * v1 + R[5] * v2) * inv;
    T[3 * (size_t)i + 2] = (R[6] * v0 + R[7] * v1 + R[8] * v2) * inv;
}

__global__ __launch_bounds__(kThreads) void k_output(int npts, const float* __restrict__ Y,
    const float* __restrict__ frame, double scale, double* __restrict__ z)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= npts) return;
    const float* R = frame + 9 * (size_t)i;
    const double y0 = Y[3 * (size_t)i], y1 = Y[3 * (size_t)i + 1], y2 = Y[3 * (size_t)i + 2];
    z[3 * (size_t)i + 0] = scale * (R[0] * y0 + R[3] * y1 + R[6] * y2);
    z[3 * (size_t)i + 1] = scale * (R[1] * y0 + R[4] * y1 + R[7] * y2);
    z[3 * (size_t)i + 2] = scale * (R[2] * y0 + R[5] * y1 + R[8] * y2);
}

__global__ __launch_bounds__(kThreads) void k_loss(long count, const float* __restrict__ Y,
    const float* __restrict__ T, float* __restrict__ dY, float* __restrict__ partials)
{
    __shared__ float red[kThreads];
    float            acc = 0.f;
    const float      g = 2.f / (float)count;
    for (long q = (long)blockIdx.x * kThreads + threadIdx.x; q < count; q += (long)gridDim.x * kThreads)
    {
        const float d = Y[q] - T[q];
        dY[q] = g * d;
        acc = fmaf(d, d, acc);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = kThreads / 2; h > 0; h >>= 1)
    {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kThreads) void k_colsum(int M, int N, const float* __restrict__ dZ,
    float* __restrict__ db)
{
    __shared__ float red[4][64];
    const int        c = threadIdx.x % 64, g = threadIdx.x / 64;
    const int        n = blockIdx.x * 64 + c;
    float            acc = 0.f;
    if (n < N)
        for (int m = g; m < M; m += 4) acc += dZ[(size_t)m * N + n];
    red[g][c] = acc;
    __syncthreads();
    if (g == 0 && n < N) db[n] = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
}

__global__ __launch_bounds__(kThreads) void k_adam(int n, float* __restrict__ w, const float* __restrict__ g,
    float* __restrict__ m, float* __restrict__ v, uint16_t* __restrict__ wq, float lr, float b1, float b2, float eps,
    float c1, float c2)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wq) wq[i] = f_to_bf16(wi);
}

__global__ __launch_bounds__(kThreads) void k_to_bf16(int n, const float* __restrict__ w, uint16_t* __restrict__ wq)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) wq[i] = f_to_bf16(w[i]);
}

__global__ __launch_bounds__(kThreads) void k_gather_rows(int nrows, int len, const float* __restrict__ src,
    const int* __restrict__ idx, float* __restrict__ dst)
{
    const long q = (long)blockIdx.x * kThreads + threadIdx.x;
    if (q >= (long)nrows * len) return;
    const long r = q / len, c = q - r * len;
    dst[q] = src[(size_t)idx[r] * len + c];
}

unsigned blocks_for(long n) { return static_cast<unsigned>((n + kThreads - 1) / kThreads); }

}  // namespace

namespace launch {

void mlp_features(hipStream_t s, int npts, int K, const double* r, const int* nbr, const float* frame,
    const float* rel, float inv_s, float* X)
{
    if (npts <= 0) return;
    hipLaunchKernelGGL(k_features, dim3(blocks_for((long)npts * K)), dim3(kThreads), 0, s, npts, K, r, nbr, frame,
        rel, inv_s, X);
}

void mlp_targets(hipStream_t s, int npts, const double* e, const float* frame, float inv, float* T)
{
    if (npts <= 0) return;
    hipLaunchKernelGGL(k_targets, dim3(blocks_for(npts)), dim3(kThreads), 0, s, npts, e, frame, inv, T);
}

void mlp_output(hipStream_t s, int npts, const float* Y, const float* frame, double scale, double* z)
{
    if (npts <= 0) return;
    hipLaunchKernelGGL(k_output, dim3(blocks_for(npts)), dim3(kThreads), 0, s, npts, Y, frame, scale, z);
}

void mlp_gemm(hipStream_t s, int M, int N, int K, const float* A, int lda, bool ta, const void* B, int ldb, bool tb,
    bool b_bf16, const float* bias, int act, const float* P, float* C)
{
    if (M <= 0 || N <= 0) return;
    const dim3 g((N + 63) / 64, (M + 63) / 64), t(kThreads);
#define IBLGF_GEMM(TA, TB, BF) \
    hipLaunchKernelGGL((k_gemm<TA, TB, BF>), g, t, 0, s, M, N, K, A, lda, B, ldb, bias, act, P, C)
    if (!ta && tb && b_bf16) IBLGF_GEMM(false, true, true);         // forward: X W^T
    else if (!ta && !tb && b_bf16) IBLGF_GEMM(false, false, true);  // backward data: dZ W
    else if (ta && !tb && !b_bf16) IBLGF_GEMM(true, false, false);  // weight gradient: dZ^T A
    else if (!ta && !tb && !b_bf16) IBLGF_GEMM(false, false, false);
    else
        throw std::runtime_error("unsupported gemm operand layout");
#undef IBLGF_GEMM
}

void mlp_loss(hipStream_t s, long count, const float* Y, const float* T, float* dY, float* partials)
{
    hipLaunchKernelGGL(k_loss, dim3(kLossBlocks), dim3(kThreads), 0, s, count, Y, T, dY, partials);
}

void mlp_colsum(hipStream_t s, int M, int N, const float* dZ, float* db)
{
    hipLaunchKernelGGL(k_colsum, dim3((N + 63) / 64), dim3(kThreads), 0, s, M, N, dZ, db);
}

void mlp_adam(hipStream_t s, int n, float* w, const float* g, float* m, float* v, uint16_t* wq, float lr, float b1,
    float b2, float eps, int step)
{
    const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(n)), dim3(kThreads), 0, s, n, w, g, m, v, wq, lr, b1, b2, eps, c1, c2);
}

void mlp_to_bf16(hipStream_t s, int n, const float* w, uint16_t* wq)
{
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(n)), dim3(kThreads), 0, s, n, w, wq);
}

void mlp_gather_rows(hipStream_t s, int nrows, int len, const float* src, const int* idx, float* dst)
{
    if (nrows <= 0) return;
    hipLaunchKernelGGL(k_gather_rows, dim3(blocks_for((long)nrows * len)), dim3(kThreads), 0, s, nrows, len, src,
        idx, dst);
}

}  // namespace launch
}  // namespace iblgf_amd
