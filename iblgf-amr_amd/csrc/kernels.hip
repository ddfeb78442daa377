// gfx950 kernels of the LGF fast-multipole path (FP64 plain path).
//
// Reference operations these kernels carry out on the device (IBLGF-AMR):
//   forward_dft / inverse_dft   Convolution::execute_fwrd_field / apply_backward,
//                               include/iblgf/utilities/convolution.hpp:226-304
//   spectral_mac                simd_prod_complex_add over all (source, target)
//                               pairs of Fmm::compute_influence_field,
//                               include/iblgf/fmm/fmm.hpp:988-1027, 1167-1195
//   transfer_gather / _scatter  Nli::nli_antrp_node / nli_intrp_node,
//                               include/iblgf/fmm/fmm_nli.hpp:101-247; cell_center_nli
//                               include/iblgf/interpolation/cell_center_nli_intrp.hpp:337-420, 591-674
//   buffer_exchange             halo fill, include/iblgf/domain/mpi/haloCommunicator.hpp:67-128
//   correction_stencil          cell_center_nli::add_source_correction,
//                               include/iblgf/interpolation/cell_center_nli_intrp.hpp:171-238
//   lgf_spectra                 LGF_Base::get_subblock + dft,
//                               include/iblgf/lgf/lgf.hpp:77-98, 221-262
#include "kernels.hpp"

#include <map>
#include <mutex>
#include <stdexcept>
#include <tuple>

namespace iblgf_amd {
namespace {

constexpr int kThreads = 256;

template<int N>
struct Geo
{
    static constexpr int n = N;
    static constexpr int E = N - 2;
    static constexpr int P = 2 * N - 1;
    static constexpr int H = P / 2 + 1;
    static constexpr int NF = H * P * P;
    static constexpr int N3 = N * N * N;
};

__device__ inline void twiddles(double2* W, int P)
{
    for (int m = threadIdx.x; m < P; m += blockDim.x)
    {
        double s, c;
        sincospi(2.0 * m / P, &s, &c);
        W[m] = make_double2(c, s);
    }
}

// ---------------------------------------------------------------------------
// Elementwise block kernels
// ---------------------------------------------------------------------------
template<int N>
__global__ __launch_bounds__(kThreads) void k_zero(double* f, const int* slots, int count, long cs)
{
    const int b = blockIdx.x;
    if (b >= count) return;
    double* p = f + blockIdx.y * cs + (size_t)slots[b] * Geo<N>::N3;
    for (int q = threadIdx.x; q < Geo<N>::N3; q += kThreads) p[q] = 0.0;
}

template<int N, class TD, class TS>
__device__ inline void copy_block_as(TD* d, const TS* s, int interior)
{
    for (int q = threadIdx.x; q < Geo<N>::N3; q += kThreads)
    {
        bool ring = false;
        if (interior)
        {
            const int i = q % N, j = (q / N) % N, k = q / (N * N);
            ring = i == 0 || j == 0 || k == 0 || i == N - 1 || j == N - 1 || k == N - 1;
        }
        d[q] = ring ? TD(0) : static_cast<TD>(s[q]);
    }
}

template<int N>
__device__ inline void copy_block(double* d, const double* s, int interior)
{
    for (int q = threadIdx.x; q < Geo<N>::N3; q += kThreads)
    {
        bool ring = false;
        if (interior)
        {
            const int i = q % N, j = (q / N) % N, k = q / (N * N);
            ring = i == 0 || j == 0 || k == 0 || i == N - 1 || j == N - 1 || k == N - 1;
        }
        d[q] = ring ? 0.0 : s[q];
    }
}

template<int N>
__global__ __launch_bounds__(kThreads) void k_copy(
    double* dst, const double* src, const int* slots, int count, int interior, long cs)
{
    const int b = blockIdx.x;
    if (b >= count) return;
    const size_t off = blockIdx.y * cs + (size_t)slots[b] * Geo<N>::N3;
    copy_block<N>(dst + off, src + off, interior);
}

template<int N>
__global__ __launch_bounds__(kThreads) void k_copy_table(
    FieldTable dst, FieldTable src, const int* slots, int count, int interior)
{
    const int b = blockIdx.x;
    if (b >= count) return;
    const size_t off = (size_t)slots[b] * Geo<N>::N3;
    const int    c = blockIdx.y;
    const bool   df = (dst.f32 >> c) & 1ull, sf = (src.f32 >> c) & 1ull;  // uniform per workgroup
    if (!df && !sf) copy_block<N>(dst.p[c] + off, src.p[c] + off, interior);
    else if (df && !sf)
        copy_block_as<N>(reinterpret_cast<float*>(dst.p[c]) + off, src.p[c] + off, interior);
    else if (!df && sf)
        copy_block_as<N>(dst.p[c] + off, reinterpret_cast<const float*>(src.p[c]) + off, interior);
    else
        copy_block_as<N>(reinterpret_cast<float*>(dst.p[c]) + off, reinterpret_cast<const float*>(src.p[c]) + off,
            interior);
}

template<int N>
__global__ __launch_bounds__(kThreads) void k_axpy(
    double* dst, const double* src, const int* slots, int count, double a, long cs)
{
#pragma clang fp contract(off)  // product and sum rounded separately (Operator::add)
    const int b = blockIdx.x;
    if (b >= count) return;
    const size_t off = blockIdx.y * cs + (size_t)slots[b] * Geo<N>::N3;
    for (int q = threadIdx.x; q < Geo<N>::N3; q += kThreads) dst[off + q] += src[off + q] * a;
}

// ---------------------------------------------------------------------------
// Separable 3-leg transfer: out[k][j][i] = sum Kz[k][c] Ky[j][b] Kx[i][a] in[c][b][a]
// Every leg is line-per-thread: a thread takes a whole line (N points along
// the leg's axis) into registers and writes its N outputs; the matrix row
// index is uniform across the workgroup, so matrix reads from LDS are
// broadcasts.  The x leg reads the input block from global memory, the y leg
// runs in place in the one LDS buffer, the z leg writes global memory.
// ---------------------------------------------------------------------------
// One thread per line of a leg: the n^2 = 324 lines of an 18^3 block fit one
// round of 384 threads.
constexpr int kTransferThreads = 384;

template<int N>
struct TransferLds
{
    double buf[Geo<N>::N3];
};

// out[base + o*stride] = sum_a K[o][a] v[a].  K is read from global memory at
// indices uniform across the workgroup, so the reads are scalar loads and
// the FMAs take the matrix from scalar registers.
template<int N>
__device__ inline void line_apply(const double (&v)[N], const double* __restrict__ K, double* out, int base,
    int stride)
{
#pragma unroll 2
    for (int o = 0; o < N; ++o)
    {
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < N; ++a) s = fma(K[o * N + a], v[a], s);
        out[base + o * stride] = s;
    }
}

// x and y legs: global block -> LDS buffer (x), then y in place.  `next`
// (optional): a block the workgroup reads next; once this block's loads are
// issued and consumed, each thread touches one 128-byte line of it, so its
// lines are on their way while the y and z legs run on LDS.  The touched value
// is held in `pend` and folded into `touch` only after the following block's
// loads are issued, so no wait is placed on it before then.
template<int N>
__device__ inline void transfer_xy(const double* __restrict__ in, TransferLds<N>& L, const double* __restrict__ Kx,
    const double* __restrict__ Ky, bool ring0 = false, const double* __restrict__ next = nullptr,
    double* pend = nullptr, double* touch = nullptr)
{
    static_assert(N * N <= kTransferThreads, "one line per thread");
    const int l = threadIdx.x;
    if (l < N * N)
    {
        // ring0: the input block's boundary ring is read as zero
        const int  lz = l / N, ly = l - lz * N;
        const bool zl = ring0 && (lz == 0 || ly == 0 || lz == N - 1 || ly == N - 1);
        double     v[N];
#pragma unroll
        for (int a = 0; a < N; ++a) v[a] = in[l * N + a];  // loaded unconditionally, then masked (no branches)
        if (touch) *touch += *pend;
        if (ring0)
        {
#pragma unroll
            for (int a = 0; a < N; ++a) v[a] = (zl || a == 0 || a == N - 1) ? 0.0 : v[a];
        }
        line_apply<N>(v, Kx, L.buf, l * N, 1);
    }
    constexpr int kLines = (Geo<N>::N3 * 8 + 127) / 128;
    if (next && l < kLines) *pend = next[l * 16];
    __syncthreads();
    if (l < N * N)
    {
        const int base = (l / N) * N * N + (l % N);
        double    v[N];
#pragma unroll
        for (int b = 0; b < N; ++b) v[b] = L.buf[base + b * N];
        line_apply<N>(v, Ky, L.buf, base, N);  // the thread owns this line
    }
    __syncthreads();
}

// z leg of line l = (j, i): acc[k] (+)= sum_c Kz[k][c] t2[c][j][i]
template<int N>
__device__ inline void leg_z_acc(const double* t2, const double* __restrict__ Kz, double (&acc)[N])
{
    const int l = threadIdx.x;
    double    v[N];
#pragma unroll
    for (int c = 0; c < N; ++c) v[c] = t2[c * N * N + l];
#pragma unroll
    for (int k = 0; k < N; ++k)  // fully unrolled: acc stays in registers
    {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < N; ++c) s = fma(Kz[k * N + c], v[c], s);
        acc[k] += s;
    }
}

// Side matrix of one leg.  mesh < 0: mats is [2][N*N] (by the child's bit on
// the leg's axis), the FMM's Lagrange operators.  mesh >= 0: mats is
// [2 relative positions][2][N*N], the cell-centred level transfers, with the
// relative position on each axis set by the field's mesh type
// (cell_center_nli_intrp.hpp:599-618): cell 1 on every axis; face component
// fi 0 on axis fi, 1 elsewhere; edge component fi 1 on axis fi, 0 elsewhere.
template<int N>
__device__ inline const double* side_mat(const double* mats, int mesh, int fi, int axis, int bit)
{
    constexpr int NN = N * N;
    if (mesh < 0) return mats + bit * NN;
    const int rel = mesh == 1 ? (axis == fi ? 0 : 1) : (mesh == 2 ? (axis == fi ? 1 : 0) : 1);
    return mats + (rel * 2 + bit) * NN;
}

// dst[parent] (=|+=) sum over listed children of K(child); one workgroup per
// parent, the children's contributions summed in registers in child order and
// written to (added into) the parent once, every point of it (deterministic).
template<int N>
__global__ __launch_bounds__(kTransferThreads) void k_transfer_gather(
    const double* __restrict__ src, double* __restrict__ dst, const int* __restrict__ entries, int count,
    const double* __restrict__ mats, int accumulate, int ring0, long cs, int mesh, int comp0)
{
    constexpr int N3 = Geo<N>::N3;
    __shared__ TransferLds<N> L;
    const int e = blockIdx.x;
    if (e >= count) return;
    src += blockIdx.y * cs;
    dst += blockIdx.y * cs;
    const int  fi = (comp0 + (int)blockIdx.y) % 3;
    const int* ent = entries + 9 * e;
    double     acc[N];
#pragma unroll
    for (int k = 0; k < N; ++k) acc[k] = 0.0;
    bool   first = true;
    double touch = 0.0, pend = 0.0;
    for (int c = 0; c < 8; ++c)
    {
        const int ch = ent[1 + c];
        if (ch < 0) continue;  // uniform across the workgroup
        if (!first) __syncthreads();  // previous z leg done with the buffer
        first = false;
        int nx = -1;  // the next listed child, touched ahead of its x leg
        for (int d = c + 1; d < 8 && nx < 0; ++d) nx = ent[1 + d];
        transfer_xy<N>(src + (size_t)ch * N3, L, side_mat<N>(mats, mesh, fi, 0, c & 1),
            side_mat<N>(mats, mesh, fi, 1, (c >> 1) & 1), ring0, nx >= 0 ? src + (size_t)nx * N3 : nullptr, &pend,
            &touch);
        if (threadIdx.x < N * N) leg_z_acc<N>(L.buf, side_mat<N>(mats, mesh, fi, 2, (c >> 2) & 1), acc);
    }
    asm volatile("" ::"v"(touch + pend));  // the touch loads are kept, their value unused
    if (threadIdx.x < N * N)
    {
        double* out = dst + (size_t)ent[0] * N3 + threadIdx.x;
#pragma unroll
        for (int k = 0; k < N; ++k) out[k * N * N] = accumulate ? out[k * N * N] + acc[k] : acc[k];
    }
}

// dst[entry.child] (=|+=) K(src[entry.parent])
template<int N>
__global__ __launch_bounds__(kTransferThreads) void k_transfer_scatter(
    const double* __restrict__ src, double* __restrict__ dst, const int* __restrict__ entries, int count,
    const double* __restrict__ mats, int accumulate, long cs, int mesh, int comp0)
{
    constexpr int N3 = Geo<N>::N3;
    __shared__ TransferLds<N> L;
    const int e = blockIdx.x;
    if (e >= count) return;
    src += blockIdx.y * cs;
    dst += blockIdx.y * cs;
    const int fi = (comp0 + (int)blockIdx.y) % 3;
    const int child = entries[3 * e], parent = entries[3 * e + 1], c = entries[3 * e + 2];
    transfer_xy<N>(src + (size_t)parent * N3, L, side_mat<N>(mats, mesh, fi, 0, c & 1),
        side_mat<N>(mats, mesh, fi, 1, (c >> 1) & 1));
    if (threadIdx.x < N * N)
    {
        double acc[N];
#pragma unroll
        for (int k = 0; k < N; ++k) acc[k] = 0.0;
        leg_z_acc<N>(L.buf, side_mat<N>(mats, mesh, fi, 2, (c >> 2) & 1), acc);
        double* out = dst + (size_t)child * N3 + threadIdx.x;
#pragma unroll
        for (int k = 0; k < N; ++k)
        {
            if (accumulate) out[k * N * N] += acc[k];
            else
                out[k * N * N] = acc[k];
        }
    }
}

// ---------------------------------------------------------------------------
// Halo fill and correction stencil
// ---------------------------------------------------------------------------
// Boundary-ring point r (0 <= r < N^3 - (N-2)^3) of a block: the z = 0 and
// z = N-1 planes, then per interior z layer the square's perimeter.
template<int N>
__device__ inline int ring_point(int r)
{
    constexpr int NN = N * N, PER = 4 * (N - 1);
    if (r < 2 * NN) return r < NN ? r : (N - 1) * NN + (r - NN);
    r -= 2 * NN;
    const int k = 1 + r / PER, p = r - (k - 1) * PER;
    int       i, j;
    if (p < N) i = p, j = 0;
    else if (p < 2 * N) i = p - N, j = N - 1;
    else if (p < 3 * N - 2) i = 0, j = 1 + (p - 2 * N);
    else
        i = N - 1, j = 1 + (p - 3 * N + 2);
    return k * NN + j * N + i;
}

// Ring points take the facing interior points of the face, edge or corner
// neighbour (the reads are interior points only, so blocks run concurrently).
template<int N>
__global__ __launch_bounds__(kThreads) void k_buffer_exchange(
    double* f, const int* slots, int count, const int* __restrict__ neighbor, long cs)
{
    constexpr int N3 = Geo<N>::N3, E = Geo<N>::E, RING = N3 - E * E * E;
    const int b = blockIdx.x;
    if (b >= count) return;
    f += blockIdx.y * cs;
    const int s = slots[b];
    for (int r = threadIdx.x; r < RING; r += kThreads)
    {
        const int q = ring_point<N>(r);
        const int i = q % N, j = (q / N) % N, k = q / (N * N);
        const int dx = i == 0 ? -1 : (i == N - 1 ? 1 : 0);
        const int dy = j == 0 ? -1 : (j == N - 1 ? 1 : 0);
        const int dz = k == 0 ? -1 : (k == N - 1 ? 1 : 0);
        const int nb = neighbor[27 * s + (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)];
        if (nb < 0) continue;
        f[(size_t)s * N3 + q] = f[(size_t)nb * N3 + q - (dz * E * N * N + dy * E * N + dx * E)];
    }
}

// c += (6 t - sum of the 6 neighbours) inv + om2 t on interior points; one
// thread per interior (i, j) column walks z with its z neighbours in
// registers.
template<int N>
__global__ __launch_bounds__(kThreads) void k_correction_stencil(
    const double* __restrict__ tt, double* __restrict__ cc, const int* slots, int count, double inv, double om2,
    long cs)
{
    constexpr int N3 = Geo<N>::N3, E = Geo<N>::E, NN = N * N;
    const int b = blockIdx.x;
    if (b >= count) return;
    const size_t  off = blockIdx.y * cs + (size_t)slots[b] * N3;
    const double* t = tt + off;
    double*       c = cc + off;
    for (int col = threadIdx.x; col < E * E; col += kThreads)
    {
        const int q0 = (1 + col / E) * N + 1 + col % E;  // (i, j) at z = 0
        double    below = t[q0], mid = t[q0 + NN];
#pragma unroll 4
        for (int k = 1; k <= E; ++k)
        {
            const int    q = q0 + k * NN;
            const double above = t[q + NN];
            double       v = c[q];
            v += 6.0 * mid * inv;
            v -= below * inv;
            v -= above * inv;
            v -= t[q - N] * inv;
            v -= t[q + N] * inv;
            v -= t[q + 1] * inv;
            v -= t[q - 1] * inv;
            v += om2 * mid;
            c[q] = v;
            below = mid;
            mid = above;
        }
    }
}

// ---------------------------------------------------------------------------
// Block DFT legs.  Every leg is a small dense matrix applied to lines of a
// block: a thread holds one input line in registers and produces its outputs;
// the twiddle index depends only on loop counters that are uniform across the
// workgroup, so twiddle reads from LDS are broadcasts.  The kx axis is the
// halved one; legs after the first run on GS kx-planes at a time so the
// intermediate fits in LDS.  Arithmetic and LDS intermediates are FP64; the
// forward transform's output spectrum is stored as T (double: plain FP64
// path; float: mixed path, where the source spectra that the MAC streams are
// FP32).
// ---------------------------------------------------------------------------
template<class T>
struct Cx;
template<>
struct Cx<double>
{
    using type = double2;
    static __device__ inline double2 make(double a, double b) { return make_double2(a, b); }
};
template<>
struct Cx<float>
{
    using type = float2;
    static __device__ inline float2 make(float a, float b) { return make_float2(a, b); }
};
__device__ inline double2 widen(double2 v) { return v; }
__device__ inline double2 widen(float2 v) { return make_double2(v.x, v.y); }

// One DFT workgroup per CU takes nearly all of the 160 KB LDS: at n = 18 a
// group of GS = 9 kx-planes with 768-1024 threads (12-16 waves) measured
// faster than two workgroups per CU with GS = 3, whose legs leave most
// threads idle.
template<int N>
struct DftPlan
{
    static constexpr int P = 2 * N - 1;
    static constexpr int H = P / 2 + 1;
    static constexpr int kLds = 163840 - 2048;
    // kx-planes per group: the twiddle table (P*N complex) plus, per plane,
    // one n x n and one n x P complex slab of intermediates; the largest
    // divisor of H that fits, so every group is full.
    static constexpr int kRaw = (kLds - P * N * 16) / (N * N * 16 + N * P * 16);
    static constexpr int largest_divisor(int h, int cap)
    {
        int d = cap < h ? cap : h;
        while (d > 1 && h % d) --d;
        return d;
    }
    static constexpr int GS = largest_divisor(H, kRaw);
};

// a * (c - i s)
template<class C, class T>
__device__ inline void cmac_conj(C a, C w, T& sr, T& si)
{
    sr = fma(a.x, w.x, fma(a.y, w.y, sr));
    si = fma(a.y, w.x, fma(-a.x, w.y, si));
}
// a * (c + i s)
template<class C, class T>
__device__ inline void cmac(C a, C w, T& sr, T& si)
{
    sr = fma(a.x, w.x, fma(-a.y, w.y, sr));
    si = fma(a.x, w.y, fma(a.y, w.x, si));
}

// Work split of one DFT leg over the workgroup: `lines` independent lines,
// each producing `outs` outputs; when there are fewer lines than threads the
// outputs of a line are divided into `parts` contiguous ranges of at most
// `per` outputs, item i = part * lines + line (threads of one wave share the
// part, so their twiddle reads are broadcasts).
template<int LINES, int OUTS, int THREADS>
struct LegSplit
{
    static constexpr int lines = LINES;
    static constexpr int raw = THREADS / LINES;
    static constexpr int parts = raw < 1 ? 1 : (raw > OUTS ? OUTS : raw);
    static constexpr int per = (OUTS + parts - 1) / parts;
    static constexpr int items = LINES * parts;
};

// Forward r2c of an n^3 block zero-padded to P^3 (pruned legs), in groups of
// GS kx-planes: the x leg for the group's planes (the block's n^2 x-lines are
// re-read from L2 per group), then the y leg and the z leg to global memory;
// output spec[kx][ky][kz].  Every leg keeps its outputs in registers, streams
// its inputs and is split over the workgroup's kFwdThreads threads as
// LegSplit (1024 threads measured faster than 768 here; the inverse kernel
// is the other way round).  Twiddles come from a table T[k][j] = w^(j k), so the inner loops
// carry no modular index arithmetic.  Arithmetic and LDS intermediates are
// FP64; the spectrum is stored as T (double: plain FP64 path; float: mixed
// path).
constexpr int kFwdThreads = 1024;

template<int N, class T>
__global__ __launch_bounds__(kFwdThreads) void k_forward_dft(const double* __restrict__ field,
    const double* __restrict__ field_in, int split, const int* __restrict__ slots, int count, int ncomp,
    typename Cx<T>::type* __restrict__ spec, long fcs, long scs)
{
    using G = Geo<N>;
    constexpr int P = G::P, H = G::H, GS = DftPlan<N>::GS;
    static_assert(GS >= 1 && H % GS == 0, "groups of whole kx-planes");
    using SX = LegSplit<N * N, GS, kFwdThreads>;   // x leg: lines (z, y), outputs kx in the group
    using SY = LegSplit<GS * N, P, kFwdThreads>;   // y leg: lines (g, z), outputs ky
    using SZ = LegSplit<GS * P, P, kFwdThreads>;   // z leg: lines (g, ky), outputs kz
    __shared__ double2 Tw[P * N];        // [k][j]
    __shared__ double2 X1[GS * N * N];   // [g][z][y]
    __shared__ double2 Tp[GS * N * P];   // [g][z][ky]
    for (int q = threadIdx.x; q < P * N; q += kFwdThreads)
    {
        const int k = q / N, j = q - k * N;
        double    sn, cs;
        sincospi(2.0 * ((j * k) % P) / P, &sn, &cs);
        Tw[q] = make_double2(cs, sn);
    }
    __syncthreads();
    // (block, component) items strided over the grid: every kx group ends
    // with a barrier, so the LDS intermediates are free for the next item.
    for (int item = blockIdx.x; item < count * ncomp; item += gridDim.x)
    {
        const int     c = item / count, b = item - c * count, slot = slots[b];
        const bool    ring0 = b >= split;  // read field_in, its boundary ring taken as zero
        const double* in = (ring0 ? field_in : field) + c * fcs + (size_t)slot * G::N3;
        typename Cx<T>::type* out = spec + c * scs + (size_t)slot * G::NF;
        for (int kx0 = 0; kx0 < H; kx0 += GS)
        {
            // x leg: real line (z, y) of N points -> planes g in [g0, g0 + per)
            for (int it = threadIdx.x; it < SX::items; it += kFwdThreads)
            {
                const int     part = it / SX::lines, line = it - part * SX::lines;
                const int     g0 = part * SX::per;
                const double* src = in + line * N;
                const int     lz = line / N, ly = line - lz * N;
                const bool    skip = ring0 && (lz == 0 || ly == 0 || lz == N - 1 || ly == N - 1);
                const int     xa = ring0 ? 1 : 0, xb = skip ? xa : (ring0 ? N - 1 : N);
                double        sr[SX::per], si[SX::per];
#pragma unroll
                for (int o = 0; o < SX::per; ++o) sr[o] = si[o] = 0.0;
#pragma unroll 2
                for (int x = xa; x < xb; ++x)
                {
                    const double v = src[x];
#pragma unroll
                    for (int o = 0; o < SX::per; ++o)
                        if (g0 + o < GS)
                        {
                            const double2 w = Tw[(kx0 + g0 + o) * N + x];
                            sr[o] = fma(v, w.x, sr[o]);
                            si[o] = fma(-v, w.y, si[o]);
                        }
                }
#pragma unroll
                for (int o = 0; o < SX::per; ++o)
                    if (g0 + o < GS) X1[(g0 + o) * N * N + line] = make_double2(sr[o], si[o]);
            }
            __syncthreads();
            // y leg: line (g, z), N inputs y -> outputs ky in [k0, k0 + per)
            for (int it = threadIdx.x; it < SY::items; it += kFwdThreads)
            {
                const int part = it / SY::lines, line = it - part * SY::lines;
                const int g = line / N, z = line - g * N;
                const int k0 = part * SY::per;
                double    sr[SY::per], si[SY::per];
#pragma unroll
                for (int o = 0; o < SY::per; ++o) sr[o] = si[o] = 0.0;
                const double2* src = X1 + g * N * N + z * N;
#pragma unroll 1
                for (int y = 0; y < N; ++y)
                {
                    const double2 a = src[y];
#pragma unroll
                    for (int o = 0; o < SY::per; ++o)
                        if (k0 + o < P) cmac_conj(a, Tw[(k0 + o) * N + y], sr[o], si[o]);
                }
#pragma unroll
                for (int o = 0; o < SY::per; ++o)
                    if (k0 + o < P) Tp[(g * N + z) * P + k0 + o] = make_double2(sr[o], si[o]);
            }
            __syncthreads();
            // z leg: line (g, ky), N inputs z -> outputs kz in [k0, k0 + per), to global memory
            for (int it = threadIdx.x; it < SZ::items; it += kFwdThreads)
            {
                const int part = it / SZ::lines, line = it - part * SZ::lines;
                const int g = line / P, ky = line - g * P, kx = kx0 + g;
                const int k0 = part * SZ::per;
                double    sr[SZ::per], si[SZ::per];
#pragma unroll
                for (int o = 0; o < SZ::per; ++o) sr[o] = si[o] = 0.0;
#pragma unroll 1
                for (int z = 0; z < N; ++z)
                {
                    const double2 a = Tp[(g * N + z) * P + ky];
#pragma unroll
                    for (int o = 0; o < SZ::per; ++o)
                        if (k0 + o < P) cmac_conj(a, Tw[(k0 + o) * N + z], sr[o], si[o]);
                }
                typename Cx<T>::type* o = out + (size_t)(kx * P + ky) * P + k0;
#pragma unroll
                for (int q = 0; q < SZ::per; ++q)
                    if (k0 + q < P) o[q] = Cx<T>::make((T)sr[q], (T)si[q]);
            }
            __syncthreads();
        }
    }
}

// Inverse c2r restricted to the output window [n-1, 2n-2]^3, scaled and
// written to (or added into) the target block, every point of it.  Per group of GS kx-planes: z leg from global
// memory, y leg in LDS, and the group's share of the c2r x leg accumulated in
// registers; every leg keeps its outputs in registers, streams its inputs and
// is split over the workgroup's kInvThreads threads as LegSplit (twice the
// forward kernel's threads: the same LDS footprint then holds 12 waves per
// CU).  Twiddles come from a table I[o][k] = w^((n-1+o) k).  FP64 arithmetic;
// the spectrum is read as T (double: plain FP64 path; float: mixed path) and
// widened on load.
constexpr int kInvThreads = 768;

template<int N, class T>
__global__ __launch_bounds__(kInvThreads) void k_inverse_dft(const typename Cx<T>::type* __restrict__ spec,
    const int* __restrict__ slots, int count, double* __restrict__ field, double scale, int accumulate, long scs,
    long fcs)
{
    using G = Geo<N>;
    constexpr int P = G::P, H = G::H, GS = DftPlan<N>::GS;
    static_assert(GS >= 1 && H % GS == 0, "groups of whole kx-planes");
    using SZ = LegSplit<GS * P, N, kInvThreads>;  // z leg: lines (g, ky), outputs z
    using SY = LegSplit<GS * N, N, kInvThreads>;  // y leg: lines (g, z), outputs y
    using SX = LegSplit<N * N, N, kInvThreads>;   // x leg: lines (z, y), outputs x
    static_assert(SX::items <= kInvThreads, "one x-leg item per thread");
    __shared__ double2 Iw[N * P];       // [o][k]
    __shared__ double2 U[GS * P * N];   // [g][ky][z]
    __shared__ double2 V[GS * N * N];   // [g][z][y]
    const int b = blockIdx.x;
    if (b >= count) return;
    const int      slot = slots[b];
    const typename Cx<T>::type* in = spec + blockIdx.y * scs + (size_t)slot * G::NF;
    for (int q = threadIdx.x; q < N * P; q += kInvThreads)
    {
        const int o = q / P, k = q - o * P;
        double    sn, cs;
        sincospi(2.0 * (((N - 1 + o) * k) % P) / P, &sn, &cs);
        Iw[q] = make_double2(cs, sn);
    }
    // x-leg item of this thread: line (z, y), outputs x in [x0, x0 + per)
    const int xpart = threadIdx.x / SX::lines, xline = threadIdx.x - xpart * SX::lines;
    const int x0 = xpart * SX::per;
    double    acc[SX::per];
#pragma unroll
    for (int x = 0; x < SX::per; ++x) acc[x] = 0.0;
    __syncthreads();
#pragma unroll
    for (int kx0 = 0; kx0 < H; kx0 += GS)
    {
        // z leg: row (g, ky) of P frequencies -> window points z in [z0, z0 + per)
        for (int it = threadIdx.x; it < SZ::items; it += kInvThreads)
        {
            const int      part = it / SZ::lines, line = it - part * SZ::lines;
            const int      g = line / P, ky = line - g * P, kx = kx0 + g;
            const int      z0 = part * SZ::per;
            const typename Cx<T>::type* row = in + (size_t)(kx * P + ky) * P;
            double         sr[SZ::per], si[SZ::per];
#pragma unroll
            for (int o = 0; o < SZ::per; ++o) sr[o] = si[o] = 0.0;
#pragma unroll 7
            for (int kz = 0; kz < P; ++kz)
            {
                const double2 a = widen(row[kz]);
#pragma unroll
                for (int o = 0; o < SZ::per; ++o)
                    if (z0 + o < N) cmac(a, Iw[(z0 + o) * P + kz], sr[o], si[o]);
            }
#pragma unroll
            for (int o = 0; o < SZ::per; ++o)
                if (z0 + o < N) U[(g * P + ky) * N + z0 + o] = make_double2(sr[o], si[o]);
        }
        __syncthreads();
        // y leg: line (g, z) of P frequencies -> window points y in [y0, y0 + per)
        for (int it = threadIdx.x; it < SY::items; it += kInvThreads)
        {
            const int part = it / SY::lines, line = it - part * SY::lines;
            const int g = line / N, z = line - g * N;
            const int y0 = part * SY::per;
            double    sr[SY::per], si[SY::per];
#pragma unroll
            for (int o = 0; o < SY::per; ++o) sr[o] = si[o] = 0.0;
#pragma unroll 7
            for (int ky = 0; ky < P; ++ky)
            {
                const double2 a = U[(g * P + ky) * N + z];
#pragma unroll
                for (int o = 0; o < SY::per; ++o)
                    if (y0 + o < N) cmac(a, Iw[(y0 + o) * P + ky], sr[o], si[o]);
            }
#pragma unroll
            for (int o = 0; o < SY::per; ++o)
                if (y0 + o < N) V[(g * N + z) * N + y0 + o] = make_double2(sr[o], si[o]);
        }
        __syncthreads();
        // The last group also writes the block; a block added into has its
        // old values loaded before that group's x leg, so the loads overlap
        // it (and live only through it).
        const bool last = kx0 + GS >= H;
        double*    out = field + blockIdx.y * fcs + (size_t)slot * G::N3 + xline * N + x0;
        double     prev[SX::per];
        if (last && accumulate && threadIdx.x < SX::items)
        {
#pragma unroll
            for (int x = 0; x < SX::per; ++x) prev[x] = x0 + x < N ? out[x] : 0.0;
        }
        // the group's part of the c2r x leg: line (z, y) -> window points x in [x0, x0 + per)
        if (threadIdx.x < SX::items)
            for (int g = 0; g < GS; ++g)
            {
                const int     kx = kx0 + g;
                const double2 a = V[g * N * N + xline];
                const double  c = (kx == 0) ? 1.0 : 2.0;  // P odd: no Nyquist plane
#pragma unroll
                for (int x = 0; x < SX::per; ++x)
                    if (x0 + x < N)
                    {
                        const double2 w = Iw[(x0 + x) * P + kx];
                        acc[x] = fma(c, fma(a.x, w.x, -a.y * w.y), acc[x]);
                    }
            }
        if (last && threadIdx.x < SX::items)
        {
#pragma unroll
            for (int x = 0; x < SX::per; ++x)
                if (x0 + x < N) out[x] = accumulate ? prev[x] + scale * acc[x] : scale * acc[x];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Spectral multiply-accumulate over sibling groups.  A workgroup owns KF
// frequencies and walks every group; the kernel spectra of the 343 offsets
// for its frequencies sit in LDS.  Each thread keeps the 8 target
// accumulators of one group at one frequency.
// ---------------------------------------------------------------------------
// Target accumulators of one (group, component) at one frequency, FP64 on
// both paths.  FP64 spectra accumulate in FP64 directly.  FP32 spectra form
// every product and the partial sums over at most 8 source entries in FP32;
// each partial sum is then added into the FP64 accumulator (an FP32 running
// sum over all of a target's sources measured twice the error).  The mixed
// path stores the finished target spectrum rounded to FP32.
template<class C> struct MacAcc;
template<> struct MacAcc<double2>
{
    double2 a[8];
    __device__ void zero()
    {
#pragma unroll
        for (int t = 0; t < 8; ++t) a[t] = make_double2(0.0, 0.0);
    }
    __device__ void add(int t, double2 g, double2 S)
    {
        a[t].x = fma(g.x, S.x, fma(-g.y, S.y, a[t].x));
        a[t].y = fma(g.x, S.y, fma(g.y, S.x, a[t].y));
    }
    __device__ void flush() {}
    __device__ double2 get(int t) const { return a[t]; }
};
template<> struct MacAcc<float2>
{
    double2 a[8];
    float2  p[8];
    __device__ void zero()
    {
#pragma unroll
        for (int t = 0; t < 8; ++t)
        {
            a[t] = make_double2(0.0, 0.0);
            p[t] = make_float2(0.f, 0.f);
        }
    }
    __device__ void add(int t, float2 g, float2 S)
    {
        p[t].x = fmaf(g.x, S.x, fmaf(-g.y, S.y, p[t].x));
        p[t].y = fmaf(g.x, S.y, fmaf(g.y, S.x, p[t].y));
    }
    __device__ void flush()
    {
#pragma unroll
        for (int t = 0; t < 8; ++t)
        {
            a[t].x += (double)p[t].x;
            a[t].y += (double)p[t].y;
            p[t] = make_float2(0.f, 0.f);
        }
    }
    __device__ double2 get(int t) const { return a[t]; }
};

// Row offset of target child t = (cx, cy, cz) relative to the entry's base row.
__device__ constexpr int mac_delta(int t) { return (t & 1) + 7 * ((t >> 1) & 1) + 49 * ((t >> 2) & 1); }

// An entry acting on every target of its group: all 8 children are computed
// (children that are not targets are never stored), every kernel row is a
// constant offset from the entry's base row, so the 8 LDS reads carry
// immediate offsets.  The base ob is at least 57 = mac_delta(7) and at most
// 342, so every row read lies in the table.
template<int KF, class C>
__device__ inline void mac_full(const C* Gs, int f, int meta, C S, MacAcc<C>& acc)
{
    const C* Gb = Gs + ((meta & 511) - 57) * KF + f;
#pragma unroll
    for (int t = 0; t < 8; ++t) acc.add(t, Gb[(57 - mac_delta(t)) * KF], S);
}

// An entry with a partial target mask, branch-free: ro holds the entry's 8
// LDS byte offsets (row * 128; targets outside the mask point at the all-zero
// row), precomputed by the plan, so each target costs one add before its LDS
// read and the 8 reads issue together.
template<int KF, class C>
__device__ inline void mac_masked(const C* Gs, int fbyte, uint4 ro, C S, MacAcc<C>& acc)
{
    const char* base = reinterpret_cast<const char*>(Gs) + fbyte;
    const unsigned w[4] = {ro.x, ro.y, ro.z, ro.w};
    C              g[8];
#pragma unroll
    for (int t = 0; t < 8; ++t)
    {
        const unsigned off = (t & 1) ? (w[t >> 1] >> 16) : (w[t >> 1] & 0xffffu);
        g[t] = *reinterpret_cast<const C*>(base + off);
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) acc.add(t, g[t], S);
}

// Entries [e, end) of one (group, component) in batches of 8, the batch's
// source loads issued before their use.
template<bool FULL, int NF, int KF, class C>
__device__ inline void mac_range(const MacSrc* __restrict__ srcs, const uint4* __restrict__ rowoff,
    const C* __restrict__ sp, const C* Gs, int f, int e, int end, MacAcc<C>& acc)
{
    const int fbyte = f * (int)sizeof(C);
    // FP32 spectra: the next batch's entries are loaded while this batch
    // multiplies, so a batch waits for one global load (its sources) instead
    // of two.  (The FP64 variant has no registers to spare for it.)
    constexpr bool kAhead = sizeof(C) == 8;
    MacSrc         nxt[8];
    if (kAhead && e + 8 <= end)
    {
#pragma unroll
        for (int i = 0; i < 8; ++i) nxt[i] = srcs[e + i];
    }
    for (; e + 8 <= end; e += 8)
    {
        MacSrc s[8];
        C      S[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) s[i] = kAhead ? nxt[i] : srcs[e + i];
#pragma unroll
        for (int i = 0; i < 8; ++i) S[i] = sp[(size_t)s[i].slot * NF];
        if (kAhead && e + 16 <= end)
        {
#pragma unroll
            for (int i = 0; i < 8; ++i) nxt[i] = srcs[e + 8 + i];
        }
#pragma unroll
        for (int i = 0; i < 8; ++i)
        {
            if constexpr (FULL) mac_full<KF, C>(Gs, f, s[i].meta, S[i], acc);
            else
                mac_masked<KF, C>(Gs, fbyte, rowoff[e + i], S[i], acc);
        }
        acc.flush();
    }
    for (; e < end; ++e)
    {
        const MacSrc s0 = srcs[e];
        const C      S0 = sp[(size_t)s0.slot * NF];
        if constexpr (FULL) mac_full<KF, C>(Gs, f, s0.meta, S0, acc);
        else
            mac_masked<KF, C>(Gs, fbyte, rowoff[e], S0, acc);
    }
    acc.flush();
}

// One launch covers the sibling groups of one target level, for k components
// of the same tree.  The work is a sequence of units (frequency chunk,
// group and component), chunk-major; a chunk is KF frequencies (KF *
// sizeof(C) = 128 B per kernel row, one cache line of a source spectrum per
// group and entry).  With W workgroups, workgroup w takes the whole chunks
// r W + w and an equal share of the units of the chunks left over: for every
// chunk it touches it stages that chunk's kernel spectra in LDS and walks its
// (group, component) pairs, the components of a group on neighbouring threads
// (same entries, same kernel rows); a thread keeps the 8 target accumulators
// of one group and component at one frequency.  With W = chunk count a
// workgroup owns one chunk; a large launch uses W = the workgroups resident
// on the device at once, so no partial last round of workgroups is left.  A
// group's entries that act on every target come first
// (nfull of them) and index the table directly; the rest carry partial masks
// and read their rows through rowoff.  COMPACT launches (few offsets used,
// the small coarse-level launches) load only the rows in olist and take
// every entry through rowoff; the others load all 343 rows.
constexpr int kMacMaxThreads = 1024;

template<int N, int KF, bool COMPACT, class T>
__global__ __launch_bounds__(kMacMaxThreads) void k_spectral_mac(const MacGroup* __restrict__ groups,
    int ngroups, const MacSrc* __restrict__ srcs, const uint4* __restrict__ rowoff, const int* __restrict__ olist,
    int nused, const typename Cx<T>::type* __restrict__ ghat, const typename Cx<T>::type* __restrict__ sspec,
    typename Cx<T>::type* __restrict__ tspec, int k, long scs, long tcs)
{
    using C = typename Cx<T>::type;
    static_assert(KF * sizeof(C) == kMacRowBytes, "row offsets assume 128-byte kernel rows");
    constexpr int NF = Geo<N>::NF;
    constexpr int NCHUNK = (NF + KF - 1) / KF;
    const int     nt = blockDim.x;  // a multiple of KF, sized to the launch's group count
    const int     GPW = nt / KF;
    __shared__ C  Gs[344 * KF];  // the chunk's rows, then one all-zero row
    // Workgroups are dealt to the 8 XCDs round-robin (workgroup i on XCD
    // i mod 8); XCD x takes a contiguous range of the logical workgroup
    // indices w (a bijection of the grid onto 0..W-1), so the units of one
    // XCD are contiguous and neighbouring groups' shared source lines meet in
    // its L2.
    const int nwg = gridDim.x, xcd = blockIdx.x & 7, q = nwg >> 3, rem = nwg & 7;
    const int w = xcd * q + min(xcd, rem) + (int)(blockIdx.x >> 3);
    const int nvg = ngroups * k;
    // Whole chunks first, chunk r W + w in round r (the workgroups of a round
    // walk the same groups together); the T = NCHUNK mod W chunks left over
    // are split as units, workgroup w taking units [Ut w / W, Ut (w+1) / W)
    // of their Ut = T nvg.
    const int  R = NCHUNK / nwg, T0 = R * nwg;
    const long Ut = (long)nvg * (NCHUNK - T0);
    const long ua = Ut * w / nwg, ub = Ut * (w + 1) / nwg;
    const int  ca = T0 + (int)(ua / nvg), cb = ub > ua ? T0 + (int)((ub - 1) / nvg) : ca - 1;
    const int  va = (int)(ua - (long)(ca - T0) * nvg), vb = (int)(ub - (long)(cb - T0) * nvg);
    const int  nseg = R + (cb - ca + 1);
    const int  rows = COMPACT ? nused : 343;
    const int  f = threadIdx.x % KF;
    for (int sg = 0; sg < nseg; ++sg)
    {
        const bool tail = sg >= R;
        const int  chunk = tail ? ca + (sg - R) : sg * nwg + w;
        const int  vlo = tail && chunk == ca ? va : 0, vhi = tail && chunk == cb ? vb : nvg;
        const int  f0 = chunk * KF;
        __syncthreads();  // the previous chunk's rows are no longer read
        for (int i = threadIdx.x; i < (rows + 1) * KF; i += nt)
        {
            const int c = i / KF, fi = i - c * KF;
            const int o = COMPACT ? olist[min(c, nused - 1)] : min(c, 342);
            Gs[i] = (c < rows && f0 + fi < NF) ? ghat[(size_t)o * NF + f0 + fi] : Cx<T>::make(0, 0);
        }
        __syncthreads();
        const int fg = min(f0 + f, NF - 1);  // tail lanes compute a valid frequency, never store
        // The workgroup walks the (Morton-ordered) groups in tiles, so a
        // source shared by neighbouring groups is re-read while its line is
        // still in L2.
        for (int vg = vlo + threadIdx.x / KF; vg < vhi; vg += GPW)
        {
            const int      g = vg / k, c = vg - g * k;
            const C*       sp = sspec + c * scs + fg;
            C*             tp = tspec + c * tcs + fg;
            const MacGroup* gp = groups + g;
            MacAcc<C>       acc;
            acc.zero();
            const int e0 = gp->begin, e2 = e0 + gp->count;
            int       e1 = e0;
            if constexpr (!COMPACT)
            {
                e1 = e0 + gp->nfull;
                mac_range<true, NF, KF, C>(srcs, rowoff, sp, Gs, f, e0, e1, acc);
            }
            mac_range<false, NF, KF, C>(srcs, rowoff, sp, Gs, f, e1, e2, acc);
            if (f0 + f < NF)
            {
                // target slots read only now: not held in registers over the entries
                int ts[8];
#pragma unroll
                for (int t = 0; t < 8; ++t) ts[t] = gp->tslot[t];
#pragma unroll
                for (int t = 0; t < 8; ++t)
                    if (ts[t] >= 0)
                    {
                        const double2 v = acc.get(t);
                        tp[(size_t)ts[t] * NF] = Cx<T>::make((T)v.x, (T)v.y);
                    }
            }
        }
    }
}

// The same launch as a persistent grid claiming work units from a device
// work queue (the autotuner's mac_policy 2): a kernel of its own, so the
// static grid's code is not changed by it.
template<int N, int KF, bool COMPACT, class T>
__global__ __launch_bounds__(kMacMaxThreads) void k_spectral_mac_queue(const MacGroup* __restrict__ groups,
    int ngroups, const MacSrc* __restrict__ srcs, const uint4* __restrict__ rowoff, const int* __restrict__ olist,
    int nused, const typename Cx<T>::type* __restrict__ ghat, const typename Cx<T>::type* __restrict__ sspec,
    typename Cx<T>::type* __restrict__ tspec, int k, long scs, long tcs, int* __restrict__ queue, int upc,
    int tail)
{
    using C = typename Cx<T>::type;
    static_assert(KF * sizeof(C) == kMacRowBytes, "row offsets assume 128-byte kernel rows");
    constexpr int NF = Geo<N>::NF;
    constexpr int NCHUNK = (NF + KF - 1) / KF;
    const int     nt = blockDim.x;  // a multiple of KF, sized to the launch's group count
    const int     GPW = nt / KF;
    __shared__ C  Gs[344 * KF];  // the chunk's rows, then one all-zero row
    const int nwg = gridDim.x, xcd = blockIdx.x & 7;  // workgroup i runs on XCD i mod 8
    const int nvg = ngroups * k;
    const int rows = COMPACT ? nused : 343;
    const int f = threadIdx.x % KF;
    // Stage chunk's kernel rows (then the zero row) in LDS.
    auto stage = [&](int chunk) {
        const int f0 = chunk * KF;
        for (int i = threadIdx.x; i < (rows + 1) * KF; i += nt)
        {
            const int c = i / KF, fi = i - c * KF;
            const int o = COMPACT ? olist[min(c, nused - 1)] : min(c, 342);
            Gs[i] = (c < rows && f0 + fi < NF) ? ghat[(size_t)o * NF + f0 + fi] : Cx<T>::make(0, 0);
        }
    };
    // The (group, component) pairs vlo..vhi-1 of a staged chunk.  The
    // workgroup walks the (Morton-ordered) groups in tiles, so a source
    // shared by neighbouring groups is re-read while its line is still in L2.
    auto walk = [&](int chunk, int vlo, int vhi) {
        const int f0 = chunk * KF;
        const int fg = min(f0 + f, NF - 1);  // tail lanes compute a valid frequency, never store
        for (int vg = vlo + threadIdx.x / KF; vg < vhi; vg += GPW)
        {
            const int       g = vg / k, c = vg - g * k;
            const C*        sp = sspec + c * scs + fg;
            C*              tp = tspec + c * tcs + fg;
            const MacGroup* gp = groups + g;
            MacAcc<C>       acc;
            acc.zero();
            const int e0 = gp->begin, e2 = e0 + gp->count;
            int       e1 = e0;
            if constexpr (!COMPACT)
            {
                e1 = e0 + gp->nfull;
                mac_range<true, NF, KF, C>(srcs, rowoff, sp, Gs, f, e0, e1, acc);
            }
            mac_range<false, NF, KF, C>(srcs, rowoff, sp, Gs, f, e1, e2, acc);
            if (f0 + f < NF)
            {
                // target slots read only now: not held in registers over the entries
                int ts[8];
#pragma unroll
                for (int t = 0; t < 8; ++t) ts[t] = gp->tslot[t];
#pragma unroll
                for (int t = 0; t < 8; ++t)
                    if (ts[t] >= 0)
                    {
                        const double2 v = acc.get(t);
                        tp[(size_t)ts[t] * NF] = Cx<T>::make((T)v.x, (T)v.y);
                    }
            }
        }
    };
    // Work queue: the launch's units are claimed one at a time from a
    // device counter, so a workgroup that drew cheap units (groups with
    // short source lists, or CUs left free by the other stream's
    // kernels) takes more of them.  The first NCHUNK - tail units are
    // whole chunks (all of the launch's (group, component) pairs); each
    // of the last `tail` chunks is split into upc runs of
    // ceil(nvg / upc) consecutive pairs, so the launch ends on short
    // units.  The units are dealt to the 8 XCDs as 8 contiguous ranges,
    // one counter each (queue[0..7]): a workgroup claims from its own
    // XCD's range, neighbouring chunks staying in one L2, and then helps
    // the other XCDs in turn.  A workgroup restages LDS only when its
    // next unit is in another chunk.  The last workgroup to find every
    // range empty (counter queue[8]) resets the counters for the next
    // launch.
    __shared__ int unit;
    const int      ug = (nvg + upc - 1) / upc, nwhole = NCHUNK - tail, nunits = nwhole + tail * upc;
    int            staged = -1, r = 0;
    auto           claim = [&]() {  // thread 0 only
        for (; r < 8; ++r)
        {
            const int x = (xcd + r) & 7;
            const int lo = (int)((long)nunits * x / 8), hi = (int)((long)nunits * (x + 1) / 8);
            const int c = lo + atomicAdd(queue + x, 1);
            if (c < hi) return c;
        }
        return nunits;
    };
    for (;;)
    {
        if (threadIdx.x == 0) unit = claim();
        __syncthreads();
        const int u = unit;
        __syncthreads();  // every thread holds u before thread 0 claims again
        if (u >= nunits) break;
        const int v = u - nwhole;
        const int chunk = v < 0 ? u : nwhole + v / upc;
        const int vlo = v < 0 ? 0 : (v % upc) * ug, vhi = v < 0 ? nvg : min(nvg, vlo + ug);
        if (chunk != staged)
        {
            stage(chunk);
            __syncthreads();
            staged = chunk;
        }
        walk(chunk, vlo, vhi);
    }
    if (threadIdx.x == 0)
    {
        __threadfence();
        if (atomicAdd(queue + 8, 1) == nwg - 1)
            for (int x = 0; x < 9; ++x) atomicExch(queue + x, 0);
    }
}

__global__ __launch_bounds__(kThreads) void k_to_float2(const double2* __restrict__ in, float2* __restrict__ out, long n)
{
    const long q = (long)blockIdx.x * kThreads + threadIdx.x;
    if (q < n) out[q] = make_float2((float)in[q].x, (float)in[q].y);
}

// ---------------------------------------------------------------------------
// Kernel spectra generation: sample, then three DFT legs through HBM scratch.
// ---------------------------------------------------------------------------
__device__ double lgf_asym(int n1, int n2, int n3)
{
    const double a = (double)n1 * n1, b = (double)n2 * n2, c = (double)n3 * n3;
    const double r = sqrt(a + b + c);
    const double r6 = r * r * r * r * r * r, r7 = r6 * r, r13 = r6 * r6 * r;
    const double pi = 3.14159265358979323846;
    double       v = -1.0 / 4.0 / pi / r;
    v = v - (a * a + b * b + c * c - 3.0 * a * b - 3.0 * b * c - 3.0 * c * a) / 16.0 / pi / r7;
    if (n1 < 600 || n2 < 600 || n3 < 600)
    {
        const double a3 = a * a * a, b3 = b * b * b, c3 = c * c * c;
        double t2 = -3.0 * (23 * (a3 * a + b3 * b + c3 * c) -
                               244 * (b3 * c + c3 * b + a3 * b + b3 * a + a3 * c + c3 * a) +
                               621 * ((a * a) * (b * b) + (b * b) * (c * c) + (c * c) * (a * a)) -
                               228 * ((a * a) * b * c + a * (b * b) * c + a * b * (c * c)));
        t2 = t2 / r13 / 4.0 * (8.0 / (768.0 * pi));
        v += t2;
    }
    return v;
}

__device__ double lgf_value(const double* g, int nmax, int x, int y, int z)
{
    x = abs(x);
    y = abs(y);
    z = abs(z);
    if (x <= nmax && y <= nmax && z <= nmax)
    {
        int t;
        if (x < y) { t = x; x = y; y = t; }
        if (x < z) { t = x; x = z; z = t; }
        if (y < z) { t = y; y = z; z = t; }
        return g[(long)x * (x + 1) * (x + 2) / 6 + (long)y * (y + 1) / 2 + z];
    }
    return lgf_asym(x, y, z);
}

template<int N>
__global__ __launch_bounds__(kThreads) void k_lgf_sample(int kind, int d, const double* __restrict__ g, int nmax,
    const double* __restrict__ iv, int kmax, double* __restrict__ out)
{
    constexpr int P = Geo<N>::P, E = Geo<N>::E;
    constexpr long total = 343L * P * P * P;
    const long     q = (long)blockIdx.x * kThreads + threadIdx.x;
    if (q >= total) return;
    const int x = q % P, y = (q / P) % P, z = (q / (P * P)) % P;
    const int o = (int)(q / ((long)P * P * P));
    const int ox = o % 7 - 3, oy = (o / 7) % 7 - 3, oz = o / 49 - 3;
    const int step = 1 << d;
    // shift = t_base - s_base = -o * e; block base = shift - (n-1)
    const int gx = (-ox * E - (N - 1) + x) * step;
    const int gy = (-oy * E - (N - 1) + y) * step;
    const int gz = (-oz * E - (N - 1) + z) * step;
    double    v;
    if (kind == 0) v = lgf_value(g, nmax, gx, gy, gz);
    else
    {
        const int ax = abs(gx), ay = abs(gy), az = abs(gz);
        v = (ax <= kmax && ay <= kmax && az <= kmax) ? iv[ax] * iv[ay] * iv[az] : 0.0;
    }
    out[q] = v;
}

// x leg: in[o][z][y][x] real -> out[o][z][y][kx]
template<int N>
__global__ __launch_bounds__(kThreads) void k_leg_x(const double* __restrict__ in, double2* __restrict__ out)
{
    constexpr int  P = Geo<N>::P, H = Geo<N>::H;
    constexpr long total = 343L * P * P * H;
    const long     q = (long)blockIdx.x * kThreads + threadIdx.x;
    if (q >= total) return;
    const int     kx = q % H;
    const long    line = q / H;
    const double* src = in + line * P;
    double        sr = 0.0, si = 0.0;
    for (int x = 0; x < P; ++x)
    {
        double s, c;
        sincospi(2.0 * ((kx * x) % P) / P, &s, &c);
        sr += src[x] * c;
        si -= src[x] * s;
    }
    out[q] = make_double2(sr, si);
}

// y leg: in[o][z][y][kx] -> out[o][z][ky][kx]
template<int N>
__global__ __launch_bounds__(kThreads) void k_leg_y(const double2* __restrict__ in, double2* __restrict__ out)
{
    constexpr int  P = Geo<N>::P, H = Geo<N>::H;
    constexpr long total = 343L * P * P * H;
    const long     q = (long)blockIdx.x * kThreads + threadIdx.x;
    if (q >= total) return;
    const int      kx = q % H, ky = (q / H) % P;
    const long     oz = q / ((long)H * P);
    const double2* src = in + oz * P * H;
    double         sr = 0.0, si = 0.0;
    for (int y = 0; y < P; ++y)
    {
        double s, c;
        sincospi(2.0 * ((ky * y) % P) / P, &s, &c);
        const double2 a = src[y * H + kx];
        sr += a.x * c + a.y * s;
        si += a.y * c - a.x * s;
    }
    out[q] = make_double2(sr, si);
}

// z leg: in[o][z][ky][kx] -> ghat[o][kx][ky][kz]
template<int N>
__global__ __launch_bounds__(kThreads) void k_leg_z(const double2* __restrict__ in, double2* __restrict__ out)
{
    constexpr int  P = Geo<N>::P, H = Geo<N>::H;
    constexpr long total = 343L * P * P * H;
    const long     q = (long)blockIdx.x * kThreads + threadIdx.x;
    if (q >= total) return;
    const int      kz = q % P, ky = (q / P) % P, kx = (q / (P * P)) % H;
    const long     o = q / ((long)P * P * H);
    const double2* src = in + o * P * P * H;
    double         sr = 0.0, si = 0.0;
    for (int z = 0; z < P; ++z)
    {
        double s, c;
        sincospi(2.0 * ((kz * z) % P) / P, &s, &c);
        const double2 a = src[((long)z * P + ky) * H + kx];
        sr += a.x * c + a.y * s;
        si += a.y * c - a.x * s;
    }
    out[q] = make_double2(sr, si);
}

// ---------------------------------------------------------------------------
// Dispatch on the block size n
// ---------------------------------------------------------------------------
template<template<int> class F, typename... Args>
void dispatch_n(int n, Args&&... args)
{
    switch (n)
    {
        case 8: F<8>::run(args...); break;
        case 10: F<10>::run(args...); break;
        case 12: F<12>::run(args...); break;
        case 16: F<16>::run(args...); break;
        case 18: F<18>::run(args...); break;
        default: throw std::runtime_error("unsupported block size n (supported: 8, 10, 12, 16, 18)");
    }
}

inline unsigned grid1(long total) { return (unsigned)((total + kThreads - 1) / kThreads); }

template<int N> struct RunZero
{
    static void run(hipStream_t s, double* f, const int* slots, int count, const Batch& b)
    {
        hipLaunchKernelGGL(k_zero<N>, dim3(count, b.k), dim3(kThreads), 0, s, f, slots, count, b.fstride);
    }
};
template<int N> struct RunCopy
{
    static void run(hipStream_t s, double* d, const double* src, const int* slots, int count, int interior,
        const Batch& b)
    {
        hipLaunchKernelGGL(k_copy<N>, dim3(count, b.k), dim3(kThreads), 0, s, d, src, slots, count, interior,
            b.fstride);
    }
};
template<int N> struct RunCopyTable
{
    static void run(hipStream_t s, const FieldTable& d, const FieldTable& src, const int* slots, int count,
        int interior, int k)
    {
        hipLaunchKernelGGL(k_copy_table<N>, dim3(count, k), dim3(kThreads), 0, s, d, src, slots, count, interior);
    }
};
template<int N> struct RunAxpy
{
    static void run(hipStream_t s, double* d, const double* src, const int* slots, int count, double a,
        const Batch& b)
    {
        hipLaunchKernelGGL(k_axpy<N>, dim3(count, b.k), dim3(kThreads), 0, s, d, src, slots, count, a, b.fstride);
    }
};
template<int N> struct RunGather
{
    static void run(hipStream_t s, const double* src, double* dst, const int* e, int count, const double* m,
        int accumulate, int ring0, const Batch& b, int mesh)
    {
        hipLaunchKernelGGL(k_transfer_gather<N>, dim3(count, b.k), dim3(kTransferThreads), 0, s, src, dst, e, count,
            m, accumulate, ring0, b.fstride, mesh, b.comp0);
    }
};
template<int N> struct RunScatter
{
    static void run(hipStream_t s, const double* src, double* dst, const int* e, int count, const double* m, int acc,
        const Batch& b, int mesh)
    {
        hipLaunchKernelGGL(k_transfer_scatter<N>, dim3(count, b.k), dim3(kTransferThreads), 0, s, src, dst, e, count,
            m, acc, b.fstride, mesh, b.comp0);
    }
};
template<int N> struct RunHalo
{
    static void run(hipStream_t s, double* f, const int* slots, int count, const int* nb, const Batch& b)
    {
        hipLaunchKernelGGL(k_buffer_exchange<N>, dim3(count, b.k), dim3(kThreads), 0, s, f, slots, count, nb,
            b.fstride);
    }
};
template<int N> struct RunStencil
{
    static void run(hipStream_t s, const double* t, double* c, const int* slots, int count, double inv, double om2,
        const Batch& b)
    {
        hipLaunchKernelGGL(k_correction_stencil<N>, dim3(count, b.k), dim3(kThreads), 0, s, t, c, slots, count, inv,
            om2, b.fstride);
    }
};
// Workgroups of a launch of `items` independent work items (a loop over
// items strided by the grid inside the kernel): as many as the device holds
// at once, at most one per item.  The device's count per (device, kernel,
// workgroup size) is queried once and kept.
template<class F>
inline int resident_grid(F kernel, int threads, long items)
{
    static std::mutex                                            lock;
    static std::map<std::tuple<int, const void*, int>, long> slots_of;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return static_cast<int>(items);
    const auto key = std::make_tuple(dev, reinterpret_cast<const void*>(kernel), threads);
    long       slots = 0;
    {
        std::lock_guard<std::mutex> g(lock);
        const auto                  it = slots_of.find(key);
        if (it != slots_of.end()) slots = it->second;
    }
    if (slots == 0)
    {
        int ncu = 0, occ = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, threads, 0) != hipSuccess || ncu <= 0 ||
            occ <= 0)
            return static_cast<int>(items);
        slots = static_cast<long>(ncu) * occ;
        std::lock_guard<std::mutex> g(lock);
        slots_of[key] = slots;
    }
    return static_cast<int>(items < slots ? items : slots);
}

template<int N> struct RunFwd
{
    static void run(hipStream_t s, int prec, const double* f, const double* fin, int split, const int* slots,
        int count, void* spec, const Batch& b)
    {
        const long items = static_cast<long>(count) * b.k;
        if (prec == 0)
            hipLaunchKernelGGL((k_forward_dft<N, double>),
                dim3(b.dft_grid ? items : resident_grid(k_forward_dft<N, double>, kFwdThreads, items)),
                dim3(kFwdThreads), 0, s, f, fin, split, slots, count, b.k, (double2*)spec, b.fstride, b.sstride);
        else
            hipLaunchKernelGGL((k_forward_dft<N, float>),
                dim3(b.dft_grid ? items : resident_grid(k_forward_dft<N, float>, kFwdThreads, items)),
                dim3(kFwdThreads), 0, s, f, fin, split, slots, count, b.k, (float2*)spec, b.fstride, b.sstride);
    }
};
template<int N> struct RunInv
{
    static void run(hipStream_t s, int prec, const void* spec, const int* slots, int count, double* f, double scale,
        int accumulate, const Batch& b)
    {
        // One workgroup per (block, component): a workgroup looping over
        // blocks, as the forward transform does, measured slower here.
        if (prec == 0)
            hipLaunchKernelGGL((k_inverse_dft<N, double>), dim3(count, b.k), dim3(kInvThreads), 0, s,
                (const double2*)spec, slots, count, f, scale, accumulate, b.sstride, b.fstride);
        else
            hipLaunchKernelGGL((k_inverse_dft<N, float>), dim3(count, b.k), dim3(kInvThreads), 0, s,
                (const float2*)spec, slots, count, f, scale, accumulate, b.sstride, b.fstride);
    }
};
// Workgroup size of a MAC launch: every (group, component) pair of the launch
// in one pass when it fits (at most kMacMaxThreads), in whole waves.
inline int mac_threads(long nvg, int kf)
{
    const long want = ((nvg * kf + 63) / 64) * 64;
    return want < 128 ? 128 : (want > kMacMaxThreads ? kMacMaxThreads : (int)want);
}

// Grid of a MAC launch.  A launch of full-size workgroups whose chunk count
// exceeds the workgroups the device holds at once is spread over exactly that
// many (each takes whole chunks in rounds and an equal share of the chunks
// left over); every other launch gives each workgroup one chunk.
template<class F>
inline int mac_grid(F kernel, int nt, int nchunk, int policy)
{
    return (nt < kMacMaxThreads || policy == 1) ? nchunk : resident_grid(kernel, nt, nchunk);
}
// Runs per tail chunk of a work-queue MAC launch: at most 8, each at least
// one round of the workgroup's thread groups.
inline int mac_units_per_chunk(long nvg, int nt, int kf)
{
    const long per = nt / kf;
    const long u = (nvg + per - 1) / per;
    return u < 1 ? 1 : (u > 8 ? 8 : (int)u);
}

template<int N> struct RunMac
{
    static void run(hipStream_t s, int prec, const MacGroup* g, int ng, const MacSrc* src, const uint4* rowoff,
        const int* olist, int nused, bool compact, const void* ghat, const void* ss, void* ts, const Batch& b,
        int* queue)
    {
#define IBLGF_MAC(KF, CP, T, C)                                                                                  \
    do                                                                                                           \
    {                                                                                                            \
        const int nt = mac_threads((long)ng * b.k, KF);                                                          \
        constexpr int nch = (Geo<N>::NF + KF - 1) / KF;                                                          \
        const int upc = mac_units_per_chunk((long)ng * b.k, nt, KF);                                             \
        if (queue)                                                                                               \
        {                                                                                                        \
            const int nwg = resident_grid(k_spectral_mac_queue<N, KF, CP, T>, nt, (long)nch * upc);              \
            const int tail = nwg < nch ? nwg : nch;                                                              \
            hipLaunchKernelGGL((k_spectral_mac_queue<N, KF, CP, T>), dim3(nwg), dim3(nt), 0, s, g, ng, src,      \
                rowoff, olist, nused, (const C*)ghat, (const C*)ss, (C*)ts, b.k, b.sstride, b.sstride, queue,    \
                upc, tail);                                                                                      \
        }                                                                                                        \
        else                                                                                                     \
        {                                                                                                        \
            const int nwg = mac_grid(k_spectral_mac<N, KF, CP, T>, nt, nch, b.mac_policy);                       \
            hipLaunchKernelGGL((k_spectral_mac<N, KF, CP, T>), dim3(nwg), dim3(nt), 0, s, g, ng, src, rowoff,    \
                olist, nused, (const C*)ghat, (const C*)ss, (C*)ts, b.k, b.sstride, b.sstride);                  \
        }                                                                                                        \
    } while (0)
        if (prec == 0)
        {
            if (compact) IBLGF_MAC(8, true, double, double2);
            else
                IBLGF_MAC(8, false, double, double2);
        }
        else
        {
            if (compact) IBLGF_MAC(16, true, float, float2);
            else
                IBLGF_MAC(16, false, float, float2);
        }
#undef IBLGF_MAC
    }
};
template<int N> struct RunLgf
{
    static void run(hipStream_t s, int kind, int d, const double* g, int nmax, const double* iv, int kmax,
        double* sample, double2* a, double2* b, double2* ghat)
    {
        constexpr int P = Geo<N>::P, H = Geo<N>::H;
        hipLaunchKernelGGL(k_lgf_sample<N>, dim3(grid1(343L * P * P * P)), dim3(kThreads), 0, s, kind, d, g, nmax,
            iv, kmax, sample);
        hipLaunchKernelGGL(k_leg_x<N>, dim3(grid1(343L * P * P * H)), dim3(kThreads), 0, s, sample, a);
        hipLaunchKernelGGL(k_leg_y<N>, dim3(grid1(343L * P * P * H)), dim3(kThreads), 0, s, a, b);
        hipLaunchKernelGGL(k_leg_z<N>, dim3(grid1(343L * P * P * H)), dim3(kThreads), 0, s, b, ghat);
    }
};

}  // namespace

namespace launch {

bool supported_n(int n) { return n == 8 || n == 10 || n == 12 || n == 16 || n == 18; }

void zero_blocks(hipStream_t s, int n, double* f, const int* slots, int count, const Batch& b)
{
    if (count > 0) dispatch_n<RunZero>(n, s, f, slots, count, b);
}
void copy_blocks(hipStream_t s, int n, double* dst, const double* src, const int* slots, int count,
    bool interior_only, const Batch& b)
{
    if (count > 0) dispatch_n<RunCopy>(n, s, dst, src, slots, count, interior_only ? 1 : 0, b);
}
void copy_blocks_table(hipStream_t s, int n, const FieldTable& dst, const FieldTable& src, const int* slots,
    int count, bool interior_only, int k)
{
    if (k < 1 || k > kMaxFieldTable) throw std::invalid_argument("copy_blocks_table: component count out of range");
    if (count > 0) dispatch_n<RunCopyTable>(n, s, dst, src, slots, count, interior_only ? 1 : 0, k);
}
void axpy_blocks(hipStream_t s, int n, double* dst, const double* src, const int* slots, int count, double a,
    const Batch& b)
{
    if (count > 0) dispatch_n<RunAxpy>(n, s, dst, src, slots, count, a, b);
}
void transfer_gather(hipStream_t s, int n, const double* src, double* dst, const int* entries, int count,
    const double* mats, bool accumulate, const Batch& b, bool children_interior, int mesh)
{
    if (count > 0)
        dispatch_n<RunGather>(
            n, s, src, dst, entries, count, mats, accumulate ? 1 : 0, children_interior ? 1 : 0, b, mesh);
}
void transfer_scatter(hipStream_t s, int n, const double* src, double* dst, const int* entries, int count,
    const double* mats, bool accumulate, const Batch& b, int mesh)
{
    if (count > 0) dispatch_n<RunScatter>(n, s, src, dst, entries, count, mats, accumulate ? 1 : 0, b, mesh);
}
void buffer_exchange(hipStream_t s, int n, double* f, const int* slots, int count, const int* neighbor,
    const Batch& b)
{
    if (count > 0) dispatch_n<RunHalo>(n, s, f, slots, count, neighbor, b);
}
void correction_stencil(hipStream_t s, int n, const double* t, double* c, const int* slots, int count,
    double inv_dx2, double omega2, const Batch& b)
{
    if (count > 0) dispatch_n<RunStencil>(n, s, t, c, slots, count, inv_dx2, omega2, b);
}
void forward_dft(hipStream_t s, int n, int prec, const double* field, const int* slots, int count, void* spec,
    const Batch& b, const double* field_in, int split)
{
    if (split < 0) split = count;
    if (split > count) throw std::invalid_argument("forward_dft: split outside the block list");
    if (split < count && !field_in) throw std::invalid_argument("forward_dft: no field for the interior-read blocks");
    if (prec_mfma(prec))
    {
        forward_dft_mfma(s, n, prec, field, slots, count, static_cast<float2*>(spec), b, field_in, split);
        return;
    }
    if (count > 0) dispatch_n<RunFwd>(n, s, prec, field, field_in, split, slots, count, spec, b);
}
void inverse_dft(hipStream_t s, int n, int prec, const void* spec, const int* slots, int count, double* field,
    double scale, bool accumulate, const Batch& b)
{
    if (prec_mfma(prec))
    {
        inverse_dft_mfma(s, n, prec, static_cast<const float2*>(spec), slots, count, field, scale, accumulate, b);
        return;
    }
    if (count > 0) dispatch_n<RunInv>(n, s, prec, spec, slots, count, field, scale, accumulate ? 1 : 0, b);
}
void spectral_mac(hipStream_t s, int n, int prec, const MacGroup* groups, int ngroups, const MacSrc* srcs,
    const uint4* rowoff, const int* olist, int nused, bool compact, const void* ghat, const void* sspec,
    void* tspec, const Batch& b, int* queue)
{
    if (nused < 1 || nused > 343) throw std::runtime_error("offset count out of range");
    if (compact && nused > kMacCompactMax) throw std::runtime_error("compact launch with too many offsets");
    if (b.k < 1 || b.k > 65535) throw std::runtime_error("component count out of range");
    if (b.mac_policy == 2 && !queue) throw std::runtime_error("work-queue launch without its counters");
    if (ngroups > 0)
        dispatch_n<RunMac>(n, s, prec, groups, ngroups, srcs, rowoff, olist, nused, compact, ghat, sspec, tspec, b,
            b.mac_policy == 2 ? queue : nullptr);
}
void to_float2(hipStream_t s, const double2* in, float2* out, long n)
{
    if (n > 0) hipLaunchKernelGGL(k_to_float2, dim3(grid1(n)), dim3(kThreads), 0, s, in, out, n);
}
void lgf_spectra(hipStream_t s, int n, int kind, int d, const double* g, int nmax, const double* iv, int kmax,
    double* sample, double2* a, double2* b, double2* ghat)
{
    dispatch_n<RunLgf>(n, s, kind, d, g, nmax, iv, kmax, sample, a, b, ghat);
}

}  // namespace launch
}  // namespace iblgf_amd
