// Screened 2-D lattice Green's functions, one per spanwise Fourier mode, and
// their free-space convolution over a uniform 2-D patch of cells.
//
// For screening c > 0 (c = k dx, k the mode's spanwise wavenumber) the
// lattice Green's function of (-Lap_h + c^2) is
//   G(n, m) = 1/pi * int_0^pi cos(m x) K(x)^-n / sqrt(a^2 - 4) dx,
//   a = 4 + c^2 - 2 cos x,  K = (a + sqrt(a^2 - 4)) / 2,  n >= m >= 0,
// and the reference's kernel value is -G (lgf/helmholtz.hpp:143-153 with
// lgf/helmholtz_lookup.hpp:60-68, 146-156, 170-181: the trapezoid integral of
// (1 - cos(m x) K^-n) / (K - 1/K) minus the origin value, whose first term is
// G(0, 0)).  Both terms are evaluated here as the one integral above.
#include "capi_common.hpp"

#include <hip/hip_runtime.h>

#include <cmath>

namespace iblgf_amd {
namespace {

constexpr int    kHelmMaxModes = 256;      // modes per call (the screenings travel as a kernel argument)
constexpr long   kHelmMaxCells = 1L << 16; // nx * ny cap of the direct convolution
constexpr double kHelmMinC = 1e-4;         // smallest screening (quadrature nodes grow as 1/c)
constexpr int    kTile = 256;

struct Screenings
{
    double c[kHelmMaxModes];
};

// Trapezoid intervals on [0, pi] for screening c and lattice offsets n + m <=
// span.  The integrand is periodic and analytic in the strip |Im x| < d, d =
// acosh(1 + c^2/2) (where a^2 = 4); there |cos(m x) K^-n| <= e^{(n+m) y}, so
// the rule's error at y = d/2 is below e^{-(q - span/2) d} / c: q >= span + 45/d
// keeps it under 1e-19 relative to G(0, 0).
__host__ __device__ inline int helm_nodes(double c, int span)
{
    const double d = acosh(1.0 + 0.5 * c * c);
    const int    q = (int)ceil(span + 45.0 / d);
    return (q + 255) & ~255;
}

// table[mode][j][i] = -G_c(max(i, j), min(i, j)), 0 <= i < nx, 0 <= j < ny.
__global__ void k_helm2d_table(Screenings sc, int nmodes, int nx, int ny, double* __restrict__ table)
{
    const long per = (long)nx * ny;
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= per * nmodes) return;
    const int    mode = (int)(t / per);
    const int    r = (int)(t - mode * per);
    const int    i = r % nx, j = r / nx;
    const int    n = i > j ? i : j, m = i > j ? j : i;
    const double c = sc.c[mode];
    const int    q = helm_nodes(c, nx + ny);
    const double h = M_PI / q;
    double       acc = 0.0;
    for (int p = 0; p <= q; ++p)
    {
        const double x = p * h;
        const double sh = sin(0.5 * x);
        const double e = c * c + 4.0 * sh * sh;  // a - 2, without cancellation near x = 0
        const double root = sqrt(e * (e + 4.0)); // sqrt(a^2 - 4)
        double       f = cos(m * x) * exp(-n * log1p(0.5 * (e + root))) / root;
        if (p == 0 || p == q) f *= 0.5;
        acc += f;
    }
    table[t] = -acc / q;
}

// dst[mode][t] = scale * sum_s table[mode][|tj - sj|][|ti - si|] * src[mode][s]
// over every cell s of the patch: one thread per target cell, the sources
// staged through LDS a tile at a time; neighbouring targets read neighbouring
// table entries.
__global__ __launch_bounds__(kTile) void k_helm2d_convolve(const double* __restrict__ table, int nx, int ny,
    const double* __restrict__ src, double* __restrict__ dst, double scale)
{
    __shared__ double s[kTile];
    const int         per = nx * ny;
    const long        off = (long)blockIdx.y * per;
    const double*     tab = table + off;
    const double*     f = src + off;
    const int         t = blockIdx.x * kTile + threadIdx.x;
    const int         ti = t % nx, tj = t / nx;
    double            acc = 0.0;
    for (int base = 0; base < per; base += kTile)
    {
        const int sidx = base + threadIdx.x;
        s[threadIdx.x] = sidx < per ? f[sidx] : 0.0;
        __syncthreads();
        if (t < per)
        {
            const int cnt = per - base < kTile ? per - base : kTile;
            int       si = base % nx, sj = base / nx;
            for (int u = 0; u < cnt; ++u)
            {
                const int dx = ti > si ? ti - si : si - ti;
                const int dy = tj > sj ? tj - sj : sj - tj;
                acc += tab[dy * nx + dx] * s[u];
                if (++si == nx)
                {
                    si = 0;
                    ++sj;
                }
            }
        }
        __syncthreads();
    }
    if (t < per) dst[off + t] = scale * acc;
}

// The same sums for patches 64 to 1024 cells wide: one workgroup per target
// row and mode, each thread up to four targets of the row; per source row the
// table row |tj - sj| and the source row are staged in LDS, so every
// multiply-add reads LDS only (the source value as a broadcast).  Sources are
// summed in index order, as above.
constexpr int kRowMin = 64, kRowMax = 4 * kTile;

__global__ __launch_bounds__(kTile) void k_helm2d_convolve_rows(const double* __restrict__ table, int nx, int ny,
    const double* __restrict__ src, double* __restrict__ dst, double scale)
{
    __shared__ double trow[kRowMax];
    __shared__ double srow[kRowMax];
    const long        off = (long)blockIdx.y * nx * ny;
    const int         tj = blockIdx.x;
    const double*     tab = table + off;
    const double*     f = src + off;
    double            acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int sj = 0; sj < ny; ++sj)
    {
        const int dy = tj > sj ? tj - sj : sj - tj;
        for (int x = threadIdx.x; x < nx; x += kTile)
        {
            trow[x] = tab[(long)dy * nx + x];
            srow[x] = f[(long)sj * nx + x];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 4; ++r)
        {
            const int ti = threadIdx.x + r * kTile;
            if (ti < nx)
            {
                double a = acc[r];
                for (int si = 0; si < nx; ++si)
                {
                    const int dx = ti > si ? ti - si : si - ti;
                    a += trow[dx] * srow[si];
                }
                acc[r] = a;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
    {
        const int ti = threadIdx.x + r * kTile;
        if (ti < nx) dst[off + (long)tj * nx + ti] = scale * acc[r];
    }
}

void check_shape(const char* who, int nmodes, int nx, int ny)
{
    if (nmodes < 1 || nmodes > kHelmMaxModes)
        throw std::invalid_argument(std::string(who) + ": 1 to 256 modes per call");
    if (nx < 1 || ny < 1 || (long)nx * ny > kHelmMaxCells)
        throw std::invalid_argument(std::string(who) + ": patch must hold 1 to 65536 cells");
}

void check_launch(const char* who)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string(who) + ": " + hipGetErrorString(e));
}

}  // namespace
}  // namespace iblgf_amd

using namespace iblgf_amd;
using namespace iblgf_amd::capi;

extern "C" {

int iblgf_gpu_helm2d_table(const double* c, int nmodes, int nx, int ny, double* table, void* stream)
{
    return guarded(nullptr, [&] {
        check_shape("iblgf_gpu_helm2d_table", nmodes, nx, ny);
        if (!c || !table) throw std::invalid_argument("iblgf_gpu_helm2d_table: null pointer");
        Screenings sc{};
        for (int k = 0; k < nmodes; ++k)
        {
            if (!(c[k] >= kHelmMinC) || !std::isfinite(c[k]))
                throw std::invalid_argument("iblgf_gpu_helm2d_table: screening c must be finite and >= 1e-4");
            sc.c[k] = c[k];
        }
        const long total = (long)nx * ny * nmodes;
        hipLaunchKernelGGL(k_helm2d_table, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
            (hipStream_t)stream, sc, nmodes, nx, ny, table);
        check_launch("iblgf_gpu_helm2d_table");
    });
}

int iblgf_gpu_helm2d_convolve(
    const double* table, int nmodes, int nx, int ny, const double* src, double* dst, double scale, void* stream)
{
    return guarded(nullptr, [&] {
        check_shape("iblgf_gpu_helm2d_convolve", nmodes, nx, ny);
        if (!table || !src || !dst) throw std::invalid_argument("iblgf_gpu_helm2d_convolve: null pointer");
        if (src == dst) throw std::invalid_argument("iblgf_gpu_helm2d_convolve: source and target must differ");
        const int per = nx * ny;
        if (nx >= kRowMin && nx <= kRowMax)
            hipLaunchKernelGGL(k_helm2d_convolve_rows, dim3((unsigned)ny, (unsigned)nmodes), dim3(kTile), 0,
                (hipStream_t)stream, table, nx, ny, src, dst, scale);
        else
            hipLaunchKernelGGL(k_helm2d_convolve, dim3((unsigned)((per + kTile - 1) / kTile), (unsigned)nmodes),
                dim3(kTile), 0, (hipStream_t)stream, table, nx, ny, src, dst, scale);
        check_launch("iblgf_gpu_helm2d_convolve");
    });
}

}  // extern "C"
