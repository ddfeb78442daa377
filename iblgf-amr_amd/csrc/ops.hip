// gfx950 kernels of the block operators and the immersed-boundary transfer.
//
// Reference operations (IBLGF-AMR):
//   block_op      Operator::laplace / gradient / divergence / curl / curl_transpose,
//                 include/iblgf/operators/operators.hpp:1387-1511, 2420-2453
//   clean_faces   Operator::clean_leaf_correction_boundary (329-513) and
//                 clean_ib_region_boundary (73-187): zero two planes per side
//   ib_spread     Operator::ib_smearing (face data, 996-1015) over the points'
//                 influence nodes, LinSysSolver::smearing (linsys.hpp:350-376)
//   ib_interp     Operator::ib_projection (844-861), LinSysSolver::projection (417-445)
//   dot / axpby   LinSysSolver::dot / add (linsys.hpp:500-531)
//
// All stencils are memory-bound (a handful of flops per 8-byte load): one
// workgroup per block streams its interior with unit-stride x, neighbours
// come from the same lines through L1/L2.  Contraction into FMA is off so
// each result is the reference's sequence of roundings.
#include "ops.hpp"

#include <stdexcept>

namespace iblgf_amd {
namespace {

constexpr int kThreads = 256;

template<int N, int OP>
__global__ __launch_bounds__(kThreads) void k_block_op(const double* __restrict__ s0, const double* __restrict__ s1,
    const double* __restrict__ s2, double* __restrict__ d0, double* __restrict__ d1, double* __restrict__ d2,
    const int* __restrict__ slots, int count, double fac)
{
#pragma clang fp contract(off)
    constexpr int E = N - 2, NN = N * N, N3 = N * N * N, E3 = E * E * E;
    const int     b = blockIdx.x;
    if (b >= count) return;
    const size_t off = (size_t)slots[b] * N3;
    for (int t = threadIdx.x; t < E3; t += kThreads)
    {
        const int i = t % E + 1, j = (t / E) % E + 1, k = t / (E * E) + 1;
        const size_t q = off + (size_t)(k * NN + j * N + i);
        if constexpr (OP == OP_LAPLACE)
        {
            double v = -6.0 * s0[q];
            v = v + s0[q - NN];
            v = v + s0[q + NN];
            v = v + s0[q - N];
            v = v + s0[q + N];
            v = v + s0[q - 1];
            v = v + s0[q + 1];
            d0[q] = v * fac;
        }
        else if constexpr (OP == OP_GRADIENT)
        {
            const double c = s0[q];
            d0[q] = fac * (c - s0[q - 1]);
            d1[q] = fac * (c - s0[q - N]);
            d2[q] = fac * (c - s0[q - NN]);
        }
        else if constexpr (OP == OP_DIVERGENCE)
        {
            double v = -s0[q];
            v = v - s1[q];
            v = v - s2[q];
            v = v + s0[q + 1];
            v = v + s1[q + N];
            v = v + s2[q + NN];
            d0[q] = v * fac;
        }
        else if constexpr (OP == OP_CURL)
        {
            double a = s2[q] - s2[q - N];
            a = a - s1[q];
            a = a + s1[q - NN];
            double c = s0[q] - s0[q - NN];
            c = c - s2[q];
            c = c + s2[q - 1];
            double e = s1[q] - s1[q - 1];
            e = e - s0[q];
            e = e + s0[q - N];
            d0[q] = a * fac;
            d1[q] = c * fac;
            d2[q] = e * fac;
        }
        else  // OP_CURL_TRANSPOSE
        {
            double a = s1[q] - s1[q + NN];
            a = a + s2[q + N];
            a = a - s2[q];
            double c = s2[q] - s2[q + 1];
            c = c + s0[q + NN];
            c = c - s0[q];
            double e = s0[q] - s0[q + N];
            e = e + s1[q + 1];
            e = e - s1[q];
            d0[q] = a * fac;
            d1[q] = c * fac;
            d2[q] = e * fac;
        }
    }
}

template<int N>
__global__ __launch_bounds__(kThreads) void k_clean_faces(double* f, const int* __restrict__ entries, int count,
    int width)
{
    constexpr int N3 = N * N * N;
    const int     b = blockIdx.x;
    if (b >= count) return;
    const int    mask = entries[2 * b + 1];
    double*      blk = f + (size_t)entries[2 * b] * N3;
    for (int q = threadIdx.x; q < N3; q += kThreads)
    {
        const int i = q % N, j = (q / N) % N, k = q / (N * N);
        const bool z = ((mask & 1) && k < width) || ((mask & 32) && k >= N - width);
        const bool y = ((mask & 2) && j < width) || ((mask & 16) && j >= N - width);
        const bool x = ((mask & 4) && i < width) || ((mask & 8) && i >= N - width);
        if (x || y || z) blk[q] = 0.0;
    }
}

// Operator::maxnorm and Operator::blockRootMeanSquare (include/iblgf/
// operators/operators.hpp:1149-1167, 1351-1367) of a field of k <= 3
// components on the listed blocks: per interior point the sum of the
// components' squares in component order; per block the largest such sum
// (out[2b]) and their total (out[2b+1]).  The square roots are the host's,
// so the largest norm is the reference's bit for bit.
template<int N>
__global__ __launch_bounds__(kThreads) void k_block_norms(const double* __restrict__ c0,
    const double* __restrict__ c1, const double* __restrict__ c2, int k, const int* __restrict__ slots, int count,
    double* __restrict__ out)
{
#pragma clang fp contract(off)
    constexpr int E = N - 2, E3 = E * E * E, N3 = N * N * N;
    __shared__ double ssum[kThreads], smax[kThreads];
    const int         b = blockIdx.x;
    if (b >= count) return;
    const size_t base = (size_t)slots[b] * N3;
    double       sum = 0.0, mx = 0.0;
    for (int t = threadIdx.x; t < E3; t += kThreads)
    {
        const int    i = t % E + 1, j = (t / E) % E + 1, kk = t / (E * E) + 1;
        const size_t q = base + i + (size_t)N * (j + (size_t)N * kk);
        double       v = c0[q];
        double       tmp = v * v;
        if (k > 1)
        {
            v = c1[q];
            tmp = tmp + v * v;
        }
        if (k > 2)
        {
            v = c2[q];
            tmp = tmp + v * v;
        }
        sum += tmp;
        mx = fmax(mx, tmp);
    }
    ssum[threadIdx.x] = sum;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1)
    {
        if ((int)threadIdx.x < w)
        {
            ssum[threadIdx.x] += ssum[threadIdx.x + w];
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        out[2 * b] = smax[0];
        out[2 * b + 1] = ssum[0];
    }
}

__global__ __launch_bounds__(kThreads) void k_ib_spread(const double* __restrict__ force,
    const int* __restrict__ row_ptr, const int* __restrict__ row_dst, const int* __restrict__ col,
    const double* __restrict__ w, int nrows, double* f0, double* f1, double* f2, int add)
{
#pragma clang fp contract(off)
    const int r = blockIdx.x * kThreads + threadIdx.x;
    if (r >= nrows) return;
    const int comp = row_dst[2 * r], q = row_dst[2 * r + 1];
    double*   f = comp == 0 ? f0 : (comp == 1 ? f1 : f2);
    double    s = add ? f[q] : 0.0;
    for (int k = row_ptr[r]; k < row_ptr[r + 1]; ++k) s = s + force[col[k]] * w[k];
    f[q] = s;
}

// dst[pairs[2e+1]] = src[pairs[2e]], whole blocks (old slot -> new slot).
template<int N>
__global__ __launch_bounds__(kThreads) void k_remap(double* __restrict__ dst, const double* __restrict__ src,
    const int* __restrict__ pairs, int count)
{
    constexpr int N3 = N * N * N;
    const int     e = blockIdx.x;
    if (e >= count) return;
    const double* a = src + (size_t)pairs[2 * e] * N3;
    double*       b = dst + (size_t)pairs[2 * e + 1] * N3;
    for (int t = threadIdx.x; t < N3; t += kThreads) b[t] = a[t];
}

// force[3 i + c] = force[3 i + c] - v[c] for every point i.
__global__ __launch_bounds__(kThreads) void k_force_shift(double* __restrict__ force, int npoints, double v0,
    double v1, double v2)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= npoints) return;
    force[3 * i] = force[3 * i] - v0;
    force[3 * i + 1] = force[3 * i + 1] - v1;
    force[3 * i + 2] = force[3 * i + 2] - v2;
}

__global__ __launch_bounds__(kThreads) void k_ib_interp(const double* __restrict__ f0,
    const double* __restrict__ f1, const double* __restrict__ f2, const int* __restrict__ row_ptr,
    const int* __restrict__ node, const double* __restrict__ w, int nrows, double* __restrict__ force)
{
#pragma clang fp contract(off)
    const int r = blockIdx.x * kThreads + threadIdx.x;
    if (r >= nrows) return;
    const int     comp = r % 3;
    const double* f = comp == 0 ? f0 : (comp == 1 ? f1 : f2);
    double        s = 0.0;
    for (int k = row_ptr[r]; k < row_ptr[r + 1]; ++k) s = s + f[node[k]] * w[k];
    force[r] = s;
}

__device__ inline double block_sum(double v, double* red)
{
    red[threadIdx.x] = v;
    __syncthreads();
    for (int h = kThreads / 2; h > 0; h >>= 1)
    {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    return red[0];
}

__global__ __launch_bounds__(kThreads) void k_dot_partial(const double* __restrict__ a,
    const double* __restrict__ b, const double* __restrict__ mask, long len, double* __restrict__ partials)
{
    __shared__ double red[kThreads];
    double            v = 0.0;
    for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < len; i += (long)gridDim.x * kThreads)
        if (!mask || mask[i] != 0.0) v = fma(a[i], b[i], v);
    const double s = block_sum(v, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void k_dot_final(const double* __restrict__ partials, int np,
    double* __restrict__ out)
{
    __shared__ double red[kThreads];
    double            v = 0.0;
    for (int i = threadIdx.x; i < np; i += kThreads) v += partials[i];
    const double s = block_sum(v, red);
    if (threadIdx.x == 0) *out = s;
}

__global__ __launch_bounds__(kThreads) void k_axpby(long len, double a, const double* __restrict__ x, double b,
    double* __restrict__ y, const double* __restrict__ mask)
{
#pragma clang fp contract(off)
    const long i = (long)blockIdx.x * kThreads + threadIdx.x;
    if (i < len && (!mask || mask[i] != 0.0)) y[i] = y[i] * b + x[i] * a;
}

// out[i] = sum_j mask_j V[i][j] w[j] for the nv rows of V (row stride len):
// one workgroup per row, the same fixed reduction order on every call.
__global__ __launch_bounds__(kThreads) void k_multi_dot(const double* __restrict__ V, int nv, long len,
    const double* __restrict__ w, const double* __restrict__ mask, double* __restrict__ out)
{
    __shared__ double red[kThreads];
    const int         i = blockIdx.x;
    if (i >= nv) return;
    const double* v = V + (size_t)i * len;
    double        acc = 0.0;
    for (long j = threadIdx.x; j < len; j += kThreads)
        if (!mask || mask[j] != 0.0) acc = fma(v[j], w[j], acc);
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) out[i] = s;
}

// x += sx * sum_i b_i X[i], y += sy * sum_i b_i Y[i] with b_i = c[i] / d[i],
// on the entries where mask is nonzero, the rows summed in order.
__global__ __launch_bounds__(kThreads) void k_multi_update(double* __restrict__ x, double sx,
    const double* __restrict__ X, double* __restrict__ y, double sy, const double* __restrict__ Y,
    const double* __restrict__ c, const double* __restrict__ d, int nv, long len, const double* __restrict__ mask)
{
    const long j = (long)blockIdx.x * kThreads + threadIdx.x;
    if (j >= len || (mask && mask[j] == 0.0)) return;
    double ax = 0.0, ay = 0.0;
    for (int i = 0; i < nv; ++i)
    {
        const double b = c[i] / d[i];
        ax = fma(b, X[(size_t)i * len + j], ax);
        ay = fma(b, Y[(size_t)i * len + j], ay);
    }
    x[j] = fma(sx, ax, x[j]);
    y[j] = fma(sy, ay, y[j]);
}

__global__ __launch_bounds__(kThreads) void k_mask_zero(long len, const double* __restrict__ mask,
    double* __restrict__ y)
{
    const long i = (long)blockIdx.x * kThreads + threadIdx.x;
    if (i < len && mask[i] == 0.0) y[i] = 0.0;
}

// Operator::nonlinear (operators.hpp:2455-2500): d = 0.25 (omega x u) on the
// interior nodes, face velocity u, edge vorticity w; then every node of the
// block times scale (Ifherk::nonlinear, ifherk.hpp:1676-1686).
template<int N>
__global__ __launch_bounds__(kThreads) void k_nonlinear(const double* __restrict__ u0,
    const double* __restrict__ u1, const double* __restrict__ u2, const double* __restrict__ w0,
    const double* __restrict__ w1, const double* __restrict__ w2, double* __restrict__ d0, double* __restrict__ d1,
    double* __restrict__ d2, const int* __restrict__ slots, int count, double scale)
{
#pragma clang fp contract(off)
    constexpr int X = 1, Y = N, Z = N * N, N3 = N * N * N;
    const int     b = blockIdx.x;
    if (b >= count) return;
    const size_t off = (size_t)slots[b] * N3;
    for (int t = threadIdx.x; t < N3; t += kThreads)
    {
        const int    i = t % N, j = (t / N) % N, k = t / (N * N);
        const size_t q = off + t;
        if (i == 0 || j == 0 || k == 0 || i == N - 1 || j == N - 1 || k == N - 1)
        {
            d0[q] = d0[q] * scale;
            d1[q] = d1[q] * scale;
            d2[q] = d2[q] * scale;
            continue;
        }
        double a = w1[q] * (u2[q] + u2[q - X]);
        a = a + w1[q + Z] * (u2[q + Z] + u2[q + Z - X]);
        a = a - w2[q] * (u1[q] + u1[q - X]);
        a = a - w2[q + Y] * (u1[q + Y] + u1[q + Y - X]);
        double c = w2[q] * (u0[q] + u0[q - Y]);
        c = c + w2[q + X] * (u0[q + X] + u0[q + X - Y]);
        c = c - w0[q] * (u2[q] + u2[q - Y]);
        c = c - w0[q + Z] * (u2[q + Z] + u2[q + Z - Y]);
        double e = w0[q] * (u1[q] + u1[q - Z]);
        e = e + w0[q + Y] * (u1[q + Y] + u1[q + Y - Z]);
        e = e - w1[q] * (u0[q] + u0[q - Z]);
        e = e - w1[q + X] * (u0[q + X] + u0[q + X - Z]);
        d0[q] = (0.25 * a) * scale;
        d1[q] = (0.25 * c) * scale;
        d2[q] = (0.25 * e) * scale;
    }
}

// dst = src * a, every node of the listed blocks.
template<int N>
__global__ __launch_bounds__(kThreads) void k_scale_copy(
    double* __restrict__ dst, const double* __restrict__ src, const int* __restrict__ slots, int count, double a)
{
    constexpr int N3 = N * N * N;
    const int     b = blockIdx.x;
    if (b >= count) return;
    const size_t off = (size_t)slots[b] * N3;
    for (int t = threadIdx.x; t < N3; t += kThreads) dst[off + t] = src[off + t] * a;
}

// f = f + v, every node of the listed blocks.
template<int N>
__global__ __launch_bounds__(kThreads) void k_add_const(double* __restrict__ f, const int* __restrict__ slots,
    int count, double v)
{
    constexpr int N3 = N * N * N;
    const int     b = blockIdx.x;
    if (b >= count) return;
    const size_t off = (size_t)slots[b] * N3;
    for (int t = threadIdx.x; t < N3; t += kThreads) f[off + t] = f[off + t] + v;
}

// Operator::smooth2zero (operators.hpp:653-760) for every neighbour i set in
// the entry's 27-bit mask {slot, mask}, ascending i: each node times the
// largest ramp value over the axes on which neighbour i lies off the block
// (ramp[0][idx] towards the low side, ramp[1][idx] towards the high side).
template<int N>
__global__ __launch_bounds__(kThreads) void k_smooth_ramp(
    double* __restrict__ f, const int* __restrict__ entries, int count, const double* __restrict__ ramp)
{
#pragma clang fp contract(off)
    constexpr int N3 = N * N * N;
    const int     b = blockIdx.x;
    if (b >= count) return;
    const int mask = entries[2 * b + 1];
    double*   blk = f + (size_t)entries[2 * b] * N3;
    for (int t = threadIdx.x; t < N3; t += kThreads)
    {
        const int i = t % N, j = (t / N) % N, k = t / (N * N);
        double    v = blk[t];
        for (int g = 0; g < 27; ++g)
        {
            if (!((mask >> g) & 1)) continue;
            const int gx = g % 3, gy = (g / 3) % 3, gz = g / 9;
            double    square = 0.0;
            if (gz != 1) square = fmax(square, ramp[(gz == 0 ? 0 : N) + k]);
            if (gy != 1) square = fmax(square, ramp[(gy == 0 ? 0 : N) + j]);
            if (gx != 1) square = fmax(square, ramp[(gx == 0 ? 0 : N) + i]);
            if (gx != 1 || gy != 1 || gz != 1) v = v * square;
        }
        blk[t] = v;
    }
}

template<template<int> class F, typename... Args>
void dispatch_n(int n, Args... args)
{
    switch (n)
    {
        case 8: F<8>::run(args...); break;
        case 10: F<10>::run(args...); break;
        case 12: F<12>::run(args...); break;
        case 16: F<16>::run(args...); break;
        case 18: F<18>::run(args...); break;
        default: throw std::runtime_error("unsupported block size n = " + std::to_string(n));
    }
}

template<int N>
struct RunOp
{
    static void run(hipStream_t s, int op, const double* const* src, double* const* dst, const int* slots, int count,
        double fac)
    {
        const dim3 g(count), t(kThreads);
#define IBLGF_OP(K)                                                                                           \
    hipLaunchKernelGGL((k_block_op<N, K>), g, t, 0, s, src[0], src[1], src[2], dst[0], dst[1], dst[2], slots, \
        count, fac)
        switch (op)
        {
            case OP_LAPLACE: IBLGF_OP(OP_LAPLACE); break;
            case OP_GRADIENT: IBLGF_OP(OP_GRADIENT); break;
            case OP_DIVERGENCE: IBLGF_OP(OP_DIVERGENCE); break;
            case OP_CURL: IBLGF_OP(OP_CURL); break;
            case OP_CURL_TRANSPOSE: IBLGF_OP(OP_CURL_TRANSPOSE); break;
            default: throw std::runtime_error("unknown block operator");
        }
#undef IBLGF_OP
    }
};

template<int N>
struct RunNorms
{
    static void run(hipStream_t s, const double* const c[3], int k, const int* slots, int count, double* out)
    {
        hipLaunchKernelGGL((k_block_norms<N>), dim3(count), dim3(kThreads), 0, s, c[0], c[1], c[2], k, slots, count,
            out);
    }
};

template<int N>
struct RunClean
{
    static void run(hipStream_t s, double* f, const int* entries, int count, int width)
    {
        hipLaunchKernelGGL((k_clean_faces<N>), dim3(count), dim3(kThreads), 0, s, f, entries, count, width);
    }
};

template<int N>
struct RunRemap
{
    static void run(hipStream_t s, double* dst, const double* src, const int* pairs, int count)
    {
        hipLaunchKernelGGL((k_remap<N>), dim3(count), dim3(kThreads), 0, s, dst, src, pairs, count);
    }
};
template<int N>
struct RunNonlinear
{
    static void run(hipStream_t s, const double* const* u, const double* const* w, double* const* d, const int* slots,
        int count, double scale)
    {
        hipLaunchKernelGGL((k_nonlinear<N>), dim3(count), dim3(kThreads), 0, s, u[0], u[1], u[2], w[0], w[1], w[2],
            d[0], d[1], d[2], slots, count, scale);
    }
};
template<int N>
struct RunScaleCopy
{
    static void run(hipStream_t s, double* dst, const double* src, const int* slots, int count, double a)
    {
        hipLaunchKernelGGL((k_scale_copy<N>), dim3(count), dim3(kThreads), 0, s, dst, src, slots, count, a);
    }
};
template<int N>
struct RunAddConst
{
    static void run(hipStream_t s, double* f, const int* slots, int count, double v)
    {
        hipLaunchKernelGGL((k_add_const<N>), dim3(count), dim3(kThreads), 0, s, f, slots, count, v);
    }
};
template<int N>
struct RunSmooth
{
    static void run(hipStream_t s, double* f, const int* entries, int count, const double* ramp)
    {
        hipLaunchKernelGGL((k_smooth_ramp<N>), dim3(count), dim3(kThreads), 0, s, f, entries, count, ramp);
    }
};

}  // namespace

namespace launch {

void nonlinear(hipStream_t s, int n, const double* const u[3], const double* const w[3], double* const d[3],
    const int* slots, int count, double scale)
{
    if (count > 0) dispatch_n<RunNonlinear>(n, s, u, w, d, slots, count, scale);
}

void remap_blocks(hipStream_t s, int n, double* dst, const double* src, const int* pairs, int count)
{
    if (count > 0) dispatch_n<RunRemap>(n, s, dst, src, pairs, count);
}

void scale_copy(hipStream_t s, int n, double* dst, const double* src, const int* slots, int count, double a)
{
    if (count > 0) dispatch_n<RunScaleCopy>(n, s, dst, src, slots, count, a);
}

void add_const(hipStream_t s, int n, double* f, const int* slots, int count, double v)
{
    if (count > 0) dispatch_n<RunAddConst>(n, s, f, slots, count, v);
}

void smooth_ramp(hipStream_t s, int n, double* f, const int* entries, int count, const double* ramp)
{
    if (count > 0) dispatch_n<RunSmooth>(n, s, f, entries, count, ramp);
}

void block_op(hipStream_t s, int n, int op, const double* const src[3], double* const dst[3], const int* slots,
    int count, double fac)
{
    if (count > 0) dispatch_n<RunOp>(n, s, op, src, dst, slots, count, fac);
}

void block_norms(hipStream_t s, int n, const double* const c[3], int k, const int* slots, int count, double* out)
{
    if (k < 1 || k > 3) throw std::runtime_error("block norms of 1 to 3 components");
    if (count > 0) dispatch_n<RunNorms>(n, s, c, k, slots, count, out);
}

void clean_faces(hipStream_t s, int n, double* f, const int* entries, int count, int width)
{
    if (width < 0 || width > n) throw std::runtime_error("clean width out of range");
    if (count > 0) dispatch_n<RunClean>(n, s, f, entries, count, width);
}

void ib_spread(hipStream_t s, const double* force, const int* row_ptr, const int* row_dst, const int* col,
    const double* w, int nrows, double* const fields[3], bool add)
{
    if (nrows <= 0) return;
    hipLaunchKernelGGL(k_ib_spread, dim3((nrows + kThreads - 1) / kThreads), dim3(kThreads), 0, s, force, row_ptr,
        row_dst, col, w, nrows, fields[0], fields[1], fields[2], add ? 1 : 0);
}

void force_shift(hipStream_t s, double* force, int npoints, const double v[3])
{
    if (npoints <= 0) return;
    hipLaunchKernelGGL(k_force_shift, dim3((npoints + kThreads - 1) / kThreads), dim3(kThreads), 0, s, force,
        npoints, v[0], v[1], v[2]);
}

void ib_interp(hipStream_t s, const double* const fields[3], const int* row_ptr, const int* node, const double* w,
    int nrows, double* force)
{
    if (nrows <= 0) return;
    hipLaunchKernelGGL(k_ib_interp, dim3((nrows + kThreads - 1) / kThreads), dim3(kThreads), 0, s, fields[0],
        fields[1], fields[2], row_ptr, node, w, nrows, force);
}

void dot(hipStream_t s, const double* a, const double* b, const double* mask, long len, double* partials,
    double* out)
{
    hipLaunchKernelGGL(k_dot_partial, dim3(kDotBlocks), dim3(kThreads), 0, s, a, b, mask, len, partials);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kThreads), 0, s, partials, kDotBlocks, out);
}

void axpby(hipStream_t s, long len, double a, const double* x, double b, double* y, const double* mask)
{
    if (len <= 0) return;
    hipLaunchKernelGGL(k_axpby, dim3((unsigned)((len + kThreads - 1) / kThreads)), dim3(kThreads), 0, s, len, a, x,
        b, y, mask);
}

void multi_dot(hipStream_t s, const double* V, int nv, long len, const double* w, const double* mask, double* out)
{
    if (nv > 0) hipLaunchKernelGGL(k_multi_dot, dim3(nv), dim3(kThreads), 0, s, V, nv, len, w, mask, out);
}

void multi_update(hipStream_t s, double* x, double sx, const double* X, double* y, double sy, const double* Y,
    const double* c, const double* d, int nv, long len, const double* mask)
{
    if (nv > 0 && len > 0)
        hipLaunchKernelGGL(k_multi_update, dim3((unsigned)((len + kThreads - 1) / kThreads)), dim3(kThreads), 0, s,
            x, sx, X, y, sy, Y, c, d, nv, len, mask);
}

void mask_zero(hipStream_t s, long len, const double* mask, double* y)
{
    if (len <= 0) return;
    hipLaunchKernelGGL(k_mask_zero, dim3((unsigned)((len + kThreads - 1) / kThreads)), dim3(kThreads), 0, s, len, mask,
        y);
}

}  // namespace launch
}  // namespace iblgf_amd
