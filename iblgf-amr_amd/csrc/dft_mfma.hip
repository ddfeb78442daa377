// Block DFT legs on the gfx950 matrix cores (split-precision paths).
//
// The same zero-padded block transforms as k_forward_dft / k_inverse_dft
// (Convolution::execute_fwrd_field / apply_backward,
// include/iblgf/utilities/convolution.hpp:226-304; pruned legs of
// src/utilities/convolution.cpp:357-570), with every 1-D leg applied as a
// real GEMM on MFMA tiles of 16 x 16:
//
//   D (outputs x lines) = A (outputs x inputs) . B (inputs x lines)
//
// A is the leg's DFT matrix (rows: output index and re/im; columns: input
// index and re/im), built once per workgroup into registers; B holds the
// block's lines, one line per tile column, so a lane of the result tile holds
// four consecutive outputs (two complex values) of one line and the stores of
// a leg are contiguous.  Intermediates between legs sit in LDS as FP32
// complex values; kx-planes go GS at a time so three workgroups share a CU.
//
// Numerics (the template parameter NUM):
//   NumBf16x3  BF16 three-term split: a = a_hi + a_lo, products
//              a_hi b_hi + a_hi b_lo + a_lo b_hi, FP32 accumulation.
//   NumF16x3   FP16 three-term split of operands scaled by powers of two
//              (each block line to |v| < 2^15, the DFT matrix by 2^14), FP32
//              accumulation.
//   NumF16     FP16, one term, same scaling.
//   NumI8x4    each operand as four signed 7-bit digits of a fixed-point
//              value (each block line scaled to |v| < 2^27, the DFT matrix by
//              2^26); the digit products with s + t <= 3 are summed exactly in
//              INT32 accumulators, one per digit weight, and combined in FP64
//              (the integer-slice scheme of Ootomo, Ozaki and Yokota, "DGEMM on
//              integer matrix multiplication unit", IJHPCA 2024, with
//              round-to-nearest digits).
// Every product of a leg is exact or rounded once into FP32 before the sum,
// so the leg's error is that of the split plus the FP32 storage of its
// outputs.
#include "kernels.hpp"

#include <map>
#include <mutex>
#include <stdexcept>
#include <tuple>

namespace iblgf_amd {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16    bf16x8;
typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 f16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float       f32x4;
typedef __attribute__((__vector_size__(4 * sizeof(int)))) int           i32x4;

// Lane layout of the 16 x 16 MFMA tiles used here (checked on the device):
// lane l supplies row l % 16 of A and column l % 16 of B with the K values
// VPL (l / 16) + j, j < VPL, and receives rows 4 (l / 16) + i, i < 4, of
// column l % 16 of D.
constexpr int kMfmaThreads = 256;

// ---- numerics -------------------------------------------------------------

struct NumBf16x3
{
    static constexpr int  VPL = 8, KS = 32;  // K values per lane and per instruction
    static constexpr int  AEXP = 0;          // DFT matrix scaled by 2^AEXP
    static constexpr bool kLineScale = false;
    static constexpr int  kTop = 0;
    struct Frag
    {
        bf16x8 h, l;
    };
    using Acc = f32x4;
    template<class T> __device__ static Frag make(const T* v)
    {
        Frag f;
#pragma unroll
        for (int j = 0; j < VPL; ++j)
        {
            const float  x = (float)v[j];
            const __bf16 h = (__bf16)x;
            f.h[j] = h;
            f.l[j] = (__bf16)(x - (float)h);
        }
        return f;
    }
    __device__ static void zero(Acc& a) { a = f32x4{0.f, 0.f, 0.f, 0.f}; }
    __device__ static void mac(Acc& a, const Frag& x, const Frag& y)
    {
        a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x.l, y.h, a, 0, 0, 0);
        a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x.h, y.l, a, 0, 0, 0);
        a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x.h, y.h, a, 0, 0, 0);
    }
    __device__ static float get(const Acc& a, int i) { return a[i]; }
};

template<bool THREE>
struct NumF16T
{
    static constexpr int  VPL = 8, KS = 32;
    static constexpr int  AEXP = 14;
    static constexpr bool kLineScale = true;
    static constexpr int  kTop = 15;  // scaled lines satisfy |v| < 2^15
    struct Frag
    {
        f16x8 h, l;
    };
    using Acc = f32x4;
    template<class T> __device__ static Frag make(const T* v)
    {
        Frag f;
#pragma unroll
        for (int j = 0; j < VPL; ++j)
        {
            const float    x = (float)v[j];
            const _Float16 h = (_Float16)x;
            f.h[j] = h;
            f.l[j] = THREE ? (_Float16)(x - (float)h) : (_Float16)0.f;
        }
        return f;
    }
    __device__ static void zero(Acc& a) { a = f32x4{0.f, 0.f, 0.f, 0.f}; }
    __device__ static void mac(Acc& a, const Frag& x, const Frag& y)
    {
        if constexpr (THREE)
        {
            a = __builtin_amdgcn_mfma_f32_16x16x32_f16(x.l, y.h, a, 0, 0, 0);
            a = __builtin_amdgcn_mfma_f32_16x16x32_f16(x.h, y.l, a, 0, 0, 0);
        }
        a = __builtin_amdgcn_mfma_f32_16x16x32_f16(x.h, y.h, a, 0, 0, 0);
    }
    __device__ static float get(const Acc& a, int i) { return a[i] * 0x1p-14f; }
};
using NumF16x3 = NumF16T<true>;
using NumF16 = NumF16T<false>;

struct NumI8x4
{
    static constexpr int  VPL = 16, KS = 64;
    static constexpr int  AEXP = 26;
    static constexpr bool kLineScale = true;
    static constexpr int  kTop = 27;  // scaled lines satisfy |v| < 2^27
    // digit s of the 16 values, four per word (value j in byte j % 4 of word j / 4)
    struct Frag
    {
        i32x4 d[4];
    };
    // one accumulator per digit weight m = s + t <= 3
    struct Acc
    {
        i32x4 m[4];
    };
    template<class T> __device__ static Frag make(const T* v)
    {
        Frag f;
#pragma unroll
        for (int w = 0; w < 4; ++w)
        {
            unsigned p0 = 0, p1 = 0, p2 = 0, p3 = 0;
#pragma unroll
            for (int b = 0; b < 4; ++b)
            {
                // X = d0 2^21 + d1 2^14 + d2 2^7 + d3, every digit rounded to
                // nearest, so |d_s| <= 64
                int       X = sizeof(T) == 4 ? (int)rintf((float)v[4 * w + b]) : (int)rint((double)v[4 * w + b]);
                const int d0 = (X + (1 << 20)) >> 21;
                X -= d0 * (1 << 21);
                const int d1 = (X + (1 << 13)) >> 14;
                X -= d1 * (1 << 14);
                const int d2 = (X + (1 << 6)) >> 7;
                const int d3 = X - d2 * (1 << 7);
                p0 |= (unsigned)(d0 & 255) << (8 * b);
                p1 |= (unsigned)(d1 & 255) << (8 * b);
                p2 |= (unsigned)(d2 & 255) << (8 * b);
                p3 |= (unsigned)(d3 & 255) << (8 * b);
            }
            f.d[0][w] = (int)p0;
            f.d[1][w] = (int)p1;
            f.d[2][w] = (int)p2;
            f.d[3][w] = (int)p3;
        }
        return f;
    }
    __device__ static void zero(Acc& a)
    {
#pragma unroll
        for (int m = 0; m < 4; ++m) a.m[m] = i32x4{0, 0, 0, 0};
    }
    __device__ static void mac(Acc& a, const Frag& x, const Frag& y)
    {
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int s = 0; s <= m; ++s)
                a.m[m] = __builtin_amdgcn_mfma_i32_16x16x64_i8(x.d[s], y.d[m - s], a.m[m], 0, 0, 0);
    }
    // sum_m I_m 2^(42 - 7 m) 2^-26: the value in units of the scaled line
    __device__ static float get(const Acc& a, int i)
    {
        double r = (double)a.m[3][i] * 0x1p-5;
        r = fma((double)a.m[2][i], 0x1p2, r);
        r = fma((double)a.m[1][i], 0x1p9, r);
        r = fma((double)a.m[0][i], 0x1p16, r);
        return (float)r;
    }
};

// ---- geometry -------------------------------------------------------------

template<int N>
struct MGeo
{
    static constexpr int P = 2 * N - 1;
    static constexpr int H = N;
    static constexpr int NF = H * P * P;
    static constexpr int N3 = N * N * N;
    static constexpr int largest_divisor(int h, int cap)
    {
        int d = cap < h ? cap : h;
        while (d > 1 && h % d) --d;
        return d;
    }
    // kx-planes per group: at most 8, so a group's re/im rows fill one tile
    static constexpr int GS = largest_divisor(H, 8);
    static constexpr int NG = H / GS;
};

// Line scale: each lane holds VALS values of one line (the line is spread
// over the four lanes l, l + 16, l + 32, l + 48); returns E with
// max |v| 2^E < 2^TOP (0 for an all-zero line) and scales the values.
template<class NUM, int VALS, class T>
__device__ inline int line_scale(T* v)
{
    if constexpr (!NUM::kLineScale) return 0;
    else
    {
        float m = 0.f;
#pragma unroll
        for (int j = 0; j < VALS; ++j) m = fmaxf(m, fabsf((float)v[j]));
        m = fmaxf(m, __shfl_xor(m, 16));
        m = fmaxf(m, __shfl_xor(m, 32));
        if (!(m > 0.f)) return 0;
        int p;
        frexpf(m, &p);  // max |v| < 2^p (rounding to FP32 is monotonic)
        const int E = NUM::kTop - p;
#pragma unroll
        for (int j = 0; j < VALS; ++j)
        {
            if constexpr (sizeof(T) == 4) v[j] = ldexpf(v[j], E);
            else
                v[j] = ldexp(v[j], E);
        }
        return E;
    }
}

// D of one tile over S K-steps.
template<class NUM, int S>
__device__ inline typename NUM::Acc tile(const typename NUM::Frag (&a)[S], const typename NUM::Frag (&b)[S])
{
    typename NUM::Acc acc;
    NUM::zero(acc);
#pragma unroll
    for (int s = 0; s < S; ++s) NUM::mac(acc, a[s], b[s]);
    return acc;
}

// ---- forward r2c ------------------------------------------------------------
//
// Per group of GS kx-planes: x leg (block lines from global memory, the group's
// kx rows), y leg and z leg; the z leg's tiles are stored straight into the
// spectrum spec[kx][ky][kz] (FP32 complex).  Stage order per block:
// X(0) | Y(0) | Z(0) + X(1) | Y(1) | ... with a barrier at each bar.
template<int N, class NUM>
__global__ __launch_bounds__(kMfmaThreads) void k_forward_mfma(const double* __restrict__ field,
    const double* __restrict__ field_in, int split, const int* __restrict__ slots, int count, int ncomp,
    float2* __restrict__ spec, long fcs, long scs)
{
    using G = MGeo<N>;
    constexpr int P = G::P, GS = G::GS, NG = G::NG, VPL = NUM::VPL, KS = NUM::KS;
    constexpr int XCT = (N * N + 15) / 16;       // x leg: lines (z, y)
    constexpr int YCT = (GS * N + 15) / 16;      // y leg: lines (g, z)
    constexpr int ZCT = (GS * P + 15) / 16;      // z leg: lines (g, ky)
    constexpr int RT = (2 * P + 15) / 16;        // y/z leg row tiles: (k, re/im)
    constexpr int S = (2 * N + KS - 1) / KS;     // y/z leg K steps: (j, re/im)
    static_assert(N <= KS, "x leg in one K step");
    __shared__ float2  X1[GS * N * N];  // [g][z][y]
    __shared__ float2  Y1[GS * P * N];  // [g][ky][z]
    __shared__ double2 Tw[P];           // e^(2 pi i m / P)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, q = lane >> 4, c16 = lane & 15;
    for (int m = threadIdx.x; m < P; m += kMfmaThreads)
    {
        double sn, cs;
        sincospi(2.0 * m / P, &sn, &cs);
        Tw[m] = make_double2(cs, sn);
    }
    __syncthreads();
    // y/z legs: out(k) = sum_j in(j) e^(-2 pi i k j / P); row r = (k, re/im),
    // column c = (j, re/im)
    typename NUM::Frag A[RT][S];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int s = 0; s < S; ++s)
        {
            double    a[VPL];
            const int r = 16 * rt + c16, k = r >> 1;
#pragma unroll
            for (int j = 0; j < VPL; ++j)
            {
                const int c = KS * s + VPL * q + j, jj = c >> 1;
                double    v = 0.0;
                if (k < P && jj < N)
                {
                    const double2 w = Tw[(k * jj) % P];
                    v = (r & 1) == 0 ? ((c & 1) == 0 ? w.x : w.y) : ((c & 1) == 0 ? -w.y : w.x);
                }
                a[j] = v * (double)(1 << NUM::AEXP);
            }
            A[rt][s] = NUM::template make<double>(a);
        }

    for (int item = blockIdx.x; item < count * ncomp; item += gridDim.x)
    {
        const int     cmp = item / count, b = item - cmp * count, slot = slots[b];
        const bool    ring0 = b >= split;
        const double* in = (ring0 ? field_in : field) + cmp * fcs + (size_t)slot * G::N3;
        float2*       out = spec + cmp * scs + (size_t)slot * G::NF;

        // x leg of group gr: rows r = (kx - kx0, re/im), K = x
        auto x_stage = [&](int gr) {
            const int kx0 = gr * GS;
            double    a[VPL];
            {
                const int r = c16, g = r >> 1, kx = kx0 + g;
#pragma unroll
                for (int j = 0; j < VPL; ++j)
                {
                    const int x = VPL * q + j;
                    double    v = 0.0;
                    if (g < GS && x < N)
                    {
                        const double2 w = Tw[(kx * x) % P];
                        v = (r & 1) == 0 ? w.x : -w.y;
                    }
                    a[j] = v * (double)(1 << NUM::AEXP);
                }
            }
            const typename NUM::Frag ax[1] = {NUM::template make<double>(a)};
            for (int ct = wave; ct < XCT; ct += kMfmaThreads / 64)
            {
                const int line = 16 * ct + c16, z = line / N, y = line - z * N;
                const bool live = line < N * N && !(ring0 && (z == 0 || y == 0 || z == N - 1 || y == N - 1));
                double     v[VPL];
#pragma unroll
                for (int j = 0; j < VPL; ++j)
                {
                    const int x = VPL * q + j;
                    v[j] = (live && x < N && !(ring0 && (x == 0 || x == N - 1))) ? in[line * N + x] : 0.0;
                }
                const int                E = line_scale<NUM, VPL>(v);
                const typename NUM::Frag bx[1] = {NUM::template make<double>(v)};
                const auto               d = tile<NUM, 1>(ax, bx);
                const float              sc = ldexpf(1.f, -E);
                if (line < N * N)
                {
                    if (2 * q < GS)
                        X1[(2 * q) * N * N + z * N + y] = make_float2(NUM::get(d, 0) * sc, NUM::get(d, 1) * sc);
                    if (2 * q + 1 < GS)
                        X1[(2 * q + 1) * N * N + z * N + y] = make_float2(NUM::get(d, 2) * sc, NUM::get(d, 3) * sc);
                }
            }
        };
        // y leg: lines (g, z), K = (y, re/im), into Y1[g][ky][z]
        auto y_stage = [&]() {
            for (int ct = wave; ct < YCT; ct += kMfmaThreads / 64)
            {
                const int line = 16 * ct + c16, g = line / N, z = line - g * N;
                const bool live = line < GS * N;
                float      v[S * VPL];
#pragma unroll
                for (int s = 0; s < S; ++s)
#pragma unroll
                    for (int j = 0; j < VPL; j += 2)
                    {
                        const int yy = (KS * s + VPL * q + j) >> 1;
                        float2    u = make_float2(0.f, 0.f);
                        if (live && yy < N) u = X1[g * N * N + z * N + yy];
                        v[s * VPL + j] = u.x;
                        v[s * VPL + j + 1] = u.y;
                    }
                const int          E = line_scale<NUM, S * VPL>(v);
                typename NUM::Frag bf[S];
#pragma unroll
                for (int s = 0; s < S; ++s) bf[s] = NUM::template make<float>(v + s * VPL);
                const float sc = ldexpf(1.f, -E);
#pragma unroll
                for (int rt = 0; rt < RT; ++rt)
                {
                    const auto d = tile<NUM, S>(A[rt], bf);
                    const int  ky = 8 * rt + 2 * q;
                    if (live && ky < P)
                        Y1[(g * P + ky) * N + z] = make_float2(NUM::get(d, 0) * sc, NUM::get(d, 1) * sc);
                    if (live && ky + 1 < P)
                        Y1[(g * P + ky + 1) * N + z] = make_float2(NUM::get(d, 2) * sc, NUM::get(d, 3) * sc);
                }
            }
        };
        // z leg: lines (g, ky), K = (z, re/im), into spec[kx0 + g][ky][kz]
        auto z_stage = [&](int gr) {
            const int kx0 = gr * GS;
            for (int ct = wave; ct < ZCT; ct += kMfmaThreads / 64)
            {
                const int line = 16 * ct + c16, g = line / P, ky = line - g * P;
                const bool live = line < GS * P;
                float      v[S * VPL];
#pragma unroll
                for (int s = 0; s < S; ++s)
#pragma unroll
                    for (int j = 0; j < VPL; j += 2)
                    {
                        const int zz = (KS * s + VPL * q + j) >> 1;
                        float2    u = make_float2(0.f, 0.f);
                        if (live && zz < N) u = Y1[(g * P + ky) * N + zz];
                        v[s * VPL + j] = u.x;
                        v[s * VPL + j + 1] = u.y;
                    }
                const int          E = line_scale<NUM, S * VPL>(v);
                typename NUM::Frag bf[S];
#pragma unroll
                for (int s = 0; s < S; ++s) bf[s] = NUM::template make<float>(v + s * VPL);
                const float sc = ldexpf(1.f, -E);
                float2*     o = out + (size_t)((kx0 + g) * P + ky) * P;
#pragma unroll
                for (int rt = 0; rt < RT; ++rt)
                {
                    const auto d = tile<NUM, S>(A[rt], bf);
                    const int  kz = 8 * rt + 2 * q;
                    if (live && kz < P) o[kz] = make_float2(NUM::get(d, 0) * sc, NUM::get(d, 1) * sc);
                    if (live && kz + 1 < P) o[kz + 1] = make_float2(NUM::get(d, 2) * sc, NUM::get(d, 3) * sc);
                }
            }
        };

        x_stage(0);
        __syncthreads();
        for (int gr = 0; gr < NG; ++gr)
        {
            y_stage();
            __syncthreads();
            z_stage(gr);
            if (gr + 1 < NG) x_stage(gr + 1);
            __syncthreads();
        }
    }
}

// ---- inverse c2r, pruned to the output window -------------------------------
//
// Per group of GS kx-planes: z leg (spectrum lines from global memory), y leg,
// and the group's share of the c2r x leg, accumulated over the groups in
// registers (each wave owns fixed tiles of x-leg lines).  The last group
// writes (or adds into) every point of the target block.  Stage order:
// Z(0) | Y(0) | X(0) + Z(1) | Y(1) | ... | X(NG-1), then the store.
template<int N, class NUM>
__global__ __launch_bounds__(kMfmaThreads) void k_inverse_mfma(const float2* __restrict__ spec,
    const int* __restrict__ slots, int count, int ncomp, double* __restrict__ field, double scale, int accumulate,
    long scs, long fcs)
{
    using G = MGeo<N>;
    constexpr int P = G::P, GS = G::GS, NG = G::NG, VPL = NUM::VPL, KS = NUM::KS;
    constexpr int ZCT = (GS * P + 15) / 16;           // z leg: lines (g, ky)
    constexpr int YCT = (GS * N + 15) / 16;           // y leg: lines (g, z)
    constexpr int XCT = (N * N + 15) / 16;            // x leg: lines (z, y)
    constexpr int RT = (2 * N + 15) / 16;             // z/y leg row tiles: (window point, re/im)
    constexpr int S = (2 * P + KS - 1) / KS;          // z/y leg K steps: (k, re/im)
    constexpr int XRT = (N + 15) / 16;                // x leg row tiles: window point
    constexpr int NW = kMfmaThreads / 64;
    constexpr int XPW = (XCT + NW - 1) / NW;          // x-leg tiles per wave
    static_assert(2 * GS <= KS, "c2r leg in one K step");
    __shared__ float2  U[GS * N * P];  // [g][z][ky]
    __shared__ float2  V[N * N * GS];  // [z][y][g]
    __shared__ double2 Tw[P];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, q = lane >> 4, c16 = lane & 15;
    for (int m = threadIdx.x; m < P; m += kMfmaThreads)
    {
        double sn, cs;
        sincospi(2.0 * m / P, &sn, &cs);
        Tw[m] = make_double2(cs, sn);
    }
    __syncthreads();
    // z/y legs: out(o) = sum_k in(k) e^(2 pi i (n-1+o) k / P); row r = (o, re/im),
    // column c = (k, re/im)
    typename NUM::Frag A[RT][S];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int s = 0; s < S; ++s)
        {
            double    a[VPL];
            const int r = 16 * rt + c16, o = r >> 1;
#pragma unroll
            for (int j = 0; j < VPL; ++j)
            {
                const int c = KS * s + VPL * q + j, k = c >> 1;
                double    v = 0.0;
                if (o < N && k < P)
                {
                    const double2 w = Tw[((N - 1 + o) * k) % P];
                    v = (r & 1) == 0 ? ((c & 1) == 0 ? w.x : -w.y) : ((c & 1) == 0 ? w.y : w.x);
                }
                a[j] = v * (double)(1 << NUM::AEXP);
            }
            A[rt][s] = NUM::template make<double>(a);
        }

    for (int item = blockIdx.x; item < count * ncomp; item += gridDim.x)
    {
        const int     cmp = item / count, b = item - cmp * count, slot = slots[b];
        const float2* in = spec + cmp * scs + (size_t)slot * G::NF;
        double*       out = field + cmp * fcs + (size_t)slot * G::N3;
        float         acc[XPW][XRT][4];
#pragma unroll
        for (int t = 0; t < XPW; ++t)
#pragma unroll
            for (int rt = 0; rt < XRT; ++rt)
#pragma unroll
                for (int i = 0; i < 4; ++i) acc[t][rt][i] = 0.f;

        // z leg of group gr: lines (g, ky) of kx = kx0 + g, K = (kz, re/im), into U[g][z][ky]
        auto z_stage = [&](int gr) {
            const int kx0 = gr * GS;
            for (int ct = wave; ct < ZCT; ct += NW)
            {
                const int     line = 16 * ct + c16, g = line / P, ky = line - g * P;
                const bool    live = line < GS * P;
                const float2* row = in + (size_t)((kx0 + g) * P + ky) * P;
                float         v[S * VPL];
#pragma unroll
                for (int s = 0; s < S; ++s)
#pragma unroll
                    for (int j = 0; j < VPL; j += 2)
                    {
                        const int kz = (KS * s + VPL * q + j) >> 1;
                        float2    u = make_float2(0.f, 0.f);
                        if (live && kz < P) u = row[kz];
                        v[s * VPL + j] = u.x;
                        v[s * VPL + j + 1] = u.y;
                    }
                const int          E = line_scale<NUM, S * VPL>(v);
                typename NUM::Frag bf[S];
#pragma unroll
                for (int s = 0; s < S; ++s) bf[s] = NUM::template make<float>(v + s * VPL);
                const float sc = ldexpf(1.f, -E);
#pragma unroll
                for (int rt = 0; rt < RT; ++rt)
                {
                    const auto d = tile<NUM, S>(A[rt], bf);
                    const int  o = 8 * rt + 2 * q;
                    if (live && o < N) U[(g * N + o) * P + ky] = make_float2(NUM::get(d, 0) * sc, NUM::get(d, 1) * sc);
                    if (live && o + 1 < N)
                        U[(g * N + o + 1) * P + ky] = make_float2(NUM::get(d, 2) * sc, NUM::get(d, 3) * sc);
                }
            }
        };
        // y leg: lines (g, z), K = (ky, re/im), into V[z][y][g]
        auto y_stage = [&]() {
            for (int ct = wave; ct < YCT; ct += NW)
            {
                const int line = 16 * ct + c16, g = line / N, z = line - g * N;
                const bool live = line < GS * N;
                float      v[S * VPL];
#pragma unroll
                for (int s = 0; s < S; ++s)
#pragma unroll
                    for (int j = 0; j < VPL; j += 2)
                    {
                        const int ky = (KS * s + VPL * q + j) >> 1;
                        float2    u = make_float2(0.f, 0.f);
                        if (live && ky < P) u = U[(g * N + z) * P + ky];
                        v[s * VPL + j] = u.x;
                        v[s * VPL + j + 1] = u.y;
                    }
                const int          E = line_scale<NUM, S * VPL>(v);
                typename NUM::Frag bf[S];
#pragma unroll
                for (int s = 0; s < S; ++s) bf[s] = NUM::template make<float>(v + s * VPL);
                const float sc = ldexpf(1.f, -E);
#pragma unroll
                for (int rt = 0; rt < RT; ++rt)
                {
                    const auto d = tile<NUM, S>(A[rt], bf);
                    const int  o = 8 * rt + 2 * q;
                    if (live && o < N) V[(z * N + o) * GS + g] = make_float2(NUM::get(d, 0) * sc, NUM::get(d, 1) * sc);
                    if (live && o + 1 < N)
                        V[(z * N + o + 1) * GS + g] = make_float2(NUM::get(d, 2) * sc, NUM::get(d, 3) * sc);
                }
            }
        };
        // c2r x leg of group gr: lines (z, y), K = (kx - kx0, re/im), rows = x,
        // out(x) = sum_kx c_kx Re(V(kx) e^(2 pi i (n-1+x) kx / P)), c_0 = 1, else 2
        auto x_stage = [&](int gr) {
            const int          kx0 = gr * GS;
            typename NUM::Frag ax[XRT][1];
#pragma unroll
            for (int rt = 0; rt < XRT; ++rt)
            {
                double    a[VPL];
                const int x = 16 * rt + c16;
#pragma unroll
                for (int j = 0; j < VPL; ++j)
                {
                    const int c = VPL * q + j, g = c >> 1, kx = kx0 + g;
                    double    v = 0.0;
                    if (x < N && g < GS)
                    {
                        const double2 w = Tw[((N - 1 + x) * kx) % P];
                        const double  ck = kx == 0 ? 1.0 : 2.0;
                        v = (c & 1) == 0 ? ck * w.x : -ck * w.y;
                    }
                    a[j] = v * (double)(1 << NUM::AEXP);
                }
                ax[rt][0] = NUM::template make<double>(a);
            }
#pragma unroll
            for (int t = 0; t < XPW; ++t)
            {
                const int ct = wave + NW * t;
                if (ct >= XCT) break;  // uniform per wave
                const int line = 16 * ct + c16;
                const bool live = line < N * N;
                float      v[VPL];
#pragma unroll
                for (int j = 0; j < VPL; j += 2)
                {
                    const int g = (VPL * q + j) >> 1;
                    float2    u = make_float2(0.f, 0.f);
                    if (live && g < GS) u = V[line * GS + g];
                    v[j] = u.x;
                    v[j + 1] = u.y;
                }
                const int                E = line_scale<NUM, VPL>(v);
                const typename NUM::Frag bx[1] = {NUM::template make<float>(v)};
                const float              sc = ldexpf(1.f, -E);
#pragma unroll
                for (int rt = 0; rt < XRT; ++rt)
                {
                    const auto d = tile<NUM, 1>(ax[rt], bx);
#pragma unroll
                    for (int i = 0; i < 4; ++i) acc[t][rt][i] += NUM::get(d, i) * sc;
                }
            }
        };

        z_stage(0);
        __syncthreads();
        for (int gr = 0; gr < NG; ++gr)
        {
            y_stage();
            __syncthreads();
            x_stage(gr);
            if (gr + 1 < NG) z_stage(gr + 1);
            __syncthreads();
        }
        // store: lane holds window points x = 16 rt + 4 q + i of line (z, y)
#pragma unroll
        for (int t = 0; t < XPW; ++t)
        {
            const int ct = wave + NW * t;
            if (ct >= XCT) break;
            const int line = 16 * ct + c16;
            if (line >= N * N) continue;
            double* o = out + line * N;
#pragma unroll
            for (int rt = 0; rt < XRT; ++rt)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                {
                    const int x = 16 * rt + 4 * q + i;
                    if (x < N)
                    {
                        const double r = scale * (double)acc[t][rt][i];
                        o[x] = accumulate ? o[x] + r : r;
                    }
                }
        }
        // the next item's z leg writes U, which this item's last y leg read
        // before the barrier above; V is rewritten only after the next barrier
    }
}

// Workgroups resident on the device at once for a kernel (CUs x occupancy),
// cached per (device, kernel).
template<class F>
int resident(F kernel, long items)
{
    static std::mutex                  lock;
    static std::map<std::pair<int, const void*>, long> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return static_cast<int>(items);
    const auto key = std::make_pair(dev, reinterpret_cast<const void*>(kernel));
    long       slots = 0;
    {
        std::lock_guard<std::mutex> g(lock);
        const auto                  it = cache.find(key);
        if (it != cache.end()) slots = it->second;
    }
    if (slots == 0)
    {
        int ncu = 0, occ = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, kMfmaThreads, 0) != hipSuccess || ncu <= 0 ||
            occ <= 0)
            return static_cast<int>(items);
        slots = static_cast<long>(ncu) * occ;
        std::lock_guard<std::mutex> g(lock);
        cache[key] = slots;
    }
    return static_cast<int>(items < slots ? items : slots);
}

template<int N, class NUM>
void fwd(hipStream_t s, const double* f, const double* fin, int split, const int* slots, int count, float2* spec,
    const Batch& b)
{
    const long items = static_cast<long>(count) * b.k;
    hipLaunchKernelGGL((k_forward_mfma<N, NUM>), dim3(b.dft_grid ? items : resident(k_forward_mfma<N, NUM>, items)),
        dim3(kMfmaThreads), 0, s, f, fin, split, slots, count, b.k, spec, b.fstride, b.sstride);
}
template<int N, class NUM>
void inv(hipStream_t s, const float2* spec, const int* slots, int count, double* f, double scale, int acc,
    const Batch& b)
{
    const long items = static_cast<long>(count) * b.k;
    hipLaunchKernelGGL((k_inverse_mfma<N, NUM>), dim3(b.dft_grid ? items : resident(k_inverse_mfma<N, NUM>, items)),
        dim3(kMfmaThreads), 0, s, spec, slots, count, b.k, f, scale, acc, b.sstride, b.fstride);
}

template<int N> struct MfmaN
{
    static void forward(int prec, hipStream_t s, const double* f, const double* fin, int split, const int* slots,
        int count, float2* spec, const Batch& b)
    {
        switch (prec)
        {
            case kPrecBf16x3: fwd<N, NumBf16x3>(s, f, fin, split, slots, count, spec, b); break;
            case kPrecF16x3: fwd<N, NumF16x3>(s, f, fin, split, slots, count, spec, b); break;
            case kPrecF16: fwd<N, NumF16>(s, f, fin, split, slots, count, spec, b); break;
            case kPrecI8x4: fwd<N, NumI8x4>(s, f, fin, split, slots, count, spec, b); break;
            default: throw std::invalid_argument("forward_dft_mfma: not a matrix-core precision");
        }
    }
    static void inverse(int prec, hipStream_t s, const float2* spec, const int* slots, int count, double* f,
        double scale, int acc, const Batch& b)
    {
        switch (prec)
        {
            case kPrecBf16x3: inv<N, NumBf16x3>(s, spec, slots, count, f, scale, acc, b); break;
            case kPrecF16x3: inv<N, NumF16x3>(s, spec, slots, count, f, scale, acc, b); break;
            case kPrecF16: inv<N, NumF16>(s, spec, slots, count, f, scale, acc, b); break;
            case kPrecI8x4: inv<N, NumI8x4>(s, spec, slots, count, f, scale, acc, b); break;
            default: throw std::invalid_argument("inverse_dft_mfma: not a matrix-core precision");
        }
    }
};

}  // namespace

namespace launch {

void forward_dft_mfma(hipStream_t s, int n, int prec, const double* field, const int* slots, int count,
    float2* spec, const Batch& b, const double* field_in, int split)
{
    if (split < 0) split = count;
    if (split > count) throw std::invalid_argument("forward_dft_mfma: split outside the block list");
    if (split < count && !field_in) throw std::invalid_argument("forward_dft_mfma: no field for the interior-read blocks");
    if (count <= 0) return;
    switch (n)
    {
        case 8: MfmaN<8>::forward(prec, s, field, field_in, split, slots, count, spec, b); break;
        case 10: MfmaN<10>::forward(prec, s, field, field_in, split, slots, count, spec, b); break;
        case 12: MfmaN<12>::forward(prec, s, field, field_in, split, slots, count, spec, b); break;
        case 16: MfmaN<16>::forward(prec, s, field, field_in, split, slots, count, spec, b); break;
        case 18: MfmaN<18>::forward(prec, s, field, field_in, split, slots, count, spec, b); break;
        default: throw std::runtime_error("unsupported block size n (supported: 8, 10, 12, 16, 18)");
    }
}

void inverse_dft_mfma(hipStream_t s, int n, int prec, const float2* spec, const int* slots, int count,
    double* field, double scale, bool accumulate, const Batch& b)
{
    if (count <= 0) return;
    const int acc = accumulate ? 1 : 0;
    switch (n)
    {
        case 8: MfmaN<8>::inverse(prec, s, spec, slots, count, field, scale, acc, b); break;
        case 10: MfmaN<10>::inverse(prec, s, spec, slots, count, field, scale, acc, b); break;
        case 12: MfmaN<12>::inverse(prec, s, spec, slots, count, field, scale, acc, b); break;
        case 16: MfmaN<16>::inverse(prec, s, spec, slots, count, field, scale, acc, b); break;
        case 18: MfmaN<18>::inverse(prec, s, spec, slots, count, field, scale, acc, b); break;
        default: throw std::runtime_error("unsupported block size n (supported: 8, 10, 12, 16, 18)");
    }
}

}  // namespace launch
}  // namespace iblgf_amd
