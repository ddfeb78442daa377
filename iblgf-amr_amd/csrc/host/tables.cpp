// Kernel and transfer tables (host, FP64).  See tables.hpp.
#include "tables.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace iblgf_amd {

namespace {
constexpr double kPi = 3.14159265358979323846264338327950288;
}

// e^{-x} I_k(x), k = 0..kmax.
//  x <= 2e4 : Miller backward recurrence I_{k-1} = I_{k+1} + (2k/x) I_k,
//             normalised with e^{-x}(I_0 + 2 sum_{k>=1} I_k) = 1.
//  x >  2e4 : Hankel large-argument expansion per order.
void scaled_bessel_i(double x, int kmax, double* out)
{
    if (x <= 0.0)
    {
        out[0] = 1.0;
        for (int k = 1; k <= kmax; ++k) out[k] = 0.0;
        return;
    }
    if (x > 2.0e4)
    {
        const double pre = 1.0 / std::sqrt(2.0 * kPi * x);
        for (int k = 0; k <= kmax; ++k)
        {
            const double mu = 4.0 * k * k;
            double       term = 1.0, sum = 1.0;
            for (int j = 1; j < 80; ++j)
            {
                const double a = 2.0 * j - 1.0;
                term *= -(mu - a * a) / (8.0 * x * j);
                sum += term;
                if (std::fabs(term) < 1e-18 * std::fabs(sum)) break;
            }
            out[k] = pre * sum;
        }
        return;
    }
    const int N = kmax + static_cast<int>(x) + static_cast<int>(12.0 * std::sqrt(x)) + 40;
    double    bkp1 = 0.0, bk = 1e-280, sum = 2.0 * bk;
    for (int k = 0; k <= kmax; ++k) out[k] = 0.0;
    for (int k = N; k >= 1; --k)
    {
        const double bkm1 = bkp1 + (2.0 * k / x) * bk;
        bkp1 = bk;
        bk = bkm1;
        const int km1 = k - 1;
        if (km1 <= kmax) out[km1] = bk;
        sum += (km1 == 0 ? 1.0 : 2.0) * bk;
        if (bk > 1e250)
        {
            bk *= 1e-250;
            bkp1 *= 1e-250;
            sum *= 1e-250;
            for (int j = std::max(km1, 0); j <= kmax; ++j) out[j] *= 1e-250;
        }
    }
    const double inv = 1.0 / sum;
    for (int k = 0; k <= kmax; ++k) out[k] *= inv;
}

void LaplaceLgfTable::build(int nmax, double h, double s_lo, double t_hi)
{
    n_max = nmax;
    const long size = index(nmax, nmax, nmax) + 1;
    std::vector<double> acc(size, 0.0);
    std::vector<double> iv(nmax + 1);
    const double        s_hi = std::log(t_hi);
    const int           nodes = static_cast<int>(std::ceil((s_hi - s_lo) / h));
    // Trapezoid in s on [s_lo, s_lo + nodes*h]; the left end is negligible.
    const double s_end = s_lo + nodes * h;
    for (int q = 0; q <= nodes; ++q)
    {
        const double s = s_lo + q * h;
        const double t = std::exp(s);
        const double w = (q == nodes ? 0.5 : 1.0) * h * t;
        scaled_bessel_i(2.0 * t, nmax, iv.data());
        long id = 0;
        for (int x = 0; x <= nmax; ++x)
        {
            const double wx = w * iv[x];
            for (int y = 0; y <= x; ++y)
            {
                const double wxy = wx * iv[y];
                for (int z = 0; z <= y; ++z) acc[id++] += wxy * iv[z];
            }
        }
    }
    // Euler-Maclaurin end correction at s_end, where the integrand in s
    // behaves as g(s) ~ C e^{-s/2}: -(h^2/12) g'(b) + (h^4/720) g'''(b).
    {
        const double t = std::exp(s_end);
        scaled_bessel_i(2.0 * t, nmax, iv.data());
        const double gc = (h * h / 24.0 - h * h * h * h / 5760.0) * t;
        long id = 0;
        for (int x = 0; x <= nmax; ++x)
            for (int y = 0; y <= x; ++y)
                for (int z = 0; z <= y; ++z) acc[id++] += gc * iv[x] * iv[y] * iv[z];
    }
    // Tail beyond T = e^{s_end}: prod ive(n_i,2t) ~ (4 pi t)^{-3/2}(1 - sum(4n_i^2-1)/(16t)).
    const double T = std::exp(s_end);
    const double c0 = std::pow(4.0 * kPi, -1.5);
    g.assign(size, 0.0);
    long id = 0;
    for (int x = 0; x <= nmax; ++x)
        for (int y = 0; y <= x; ++y)
            for (int z = 0; z <= y; ++z)
            {
                const double m = (4.0 * x * x - 1.0) + (4.0 * y * y - 1.0) + (4.0 * z * z - 1.0);
                const double tail = c0 * (2.0 / std::sqrt(T) - (m / 16.0) * (2.0 / 3.0) * std::pow(T, -1.5));
                g[id] = -(acc[id] + tail);
                ++id;
            }
}

double LaplaceLgfTable::asym(int n1, int n2, int n3)
{
    // include/iblgf/lgf/lgf_gl_lookup.hpp:74-119
    const double n1_2 = (double)n1 * n1, n2_2 = (double)n2 * n2, n3_2 = (double)n3 * n3;
    const double n_abs = std::sqrt(n1_2 + n2_2 + n3_2);
    const double n1_6 = n1_2 * n1_2 * n1_2, n2_6 = n2_2 * n2_2 * n2_2, n3_6 = n3_2 * n3_2 * n3_2;
    const double n1_8 = n1_6 * n1_2, n2_8 = n2_6 * n2_2, n3_8 = n3_6 * n3_2;
    const double n_abs_6 = n_abs * n_abs * n_abs * n_abs * n_abs * n_abs;
    const double n_abs_7 = n_abs_6 * n_abs;
    const double n_abs_13 = n_abs_6 * n_abs_6 * n_abs;
    double       r = -1.0 / 4.0 / kPi / n_abs;
    r = r - (n1_2 * n1_2 + n2_2 * n2_2 + n3_2 * n3_2 - 3.0 * n1_2 * n2_2 - 3.0 * n2_2 * n3_2 -
                3.0 * n3_2 * n1_2) /
                16.0 / kPi / n_abs_7;
    if (n1 < 600 || n2 < 600 || n3 < 600)
    {
        const double coef = 8.0 / (768.0 * kPi);
        double       t2 = -3.0 * (23 * (n1_8 + n2_8 + n3_8) -
                               244 * (n2_6 * n3_2 + n3_6 * n2_2 + n1_6 * n2_2 + n2_6 * n1_2 +
                                         n1_6 * n3_2 + n3_6 * n1_2) +
                               621 * ((n1_2 * n1_2) * (n2_2 * n2_2) + (n2_2 * n2_2) * (n3_2 * n3_2) +
                                         (n3_2 * n3_2) * (n1_2 * n1_2)) -
                               228 * ((n1_2 * n1_2) * n2_2 * n3_2 + n1_2 * (n2_2 * n2_2) * n3_2 +
                                         n1_2 * n2_2 * (n3_2 * n3_2)));
        t2 = t2 / n_abs_13 / 4.0 * coef;
        r += t2;
    }
    return r;
}

double LaplaceLgfTable::get(int x, int y, int z) const
{
    x = std::abs(x);
    y = std::abs(y);
    z = std::abs(z);
    if (x <= n_max && y <= n_max && z <= n_max)
    {
        if (x < y) std::swap(x, y);
        if (x < z) std::swap(x, z);
        if (y < z) std::swap(y, z);
        return g[index(x, y, z)];
    }
    return asym(x, y, z);
}

double if_kernel(int n1, int n2, int n3, double alpha)
{
    const int kmax = std::max(std::abs(n1), std::max(std::abs(n2), std::abs(n3)));
    std::vector<double> iv(kmax + 1);
    scaled_bessel_i(2.0 * alpha, kmax, iv.data());
    return iv[std::abs(n1)] * iv[std::abs(n2)] * iv[std::abs(n3)];
}

namespace {
// Barycentric Lagrange columns for fine positions c/2 (c = 1..2n-2), then the
// child-1 columns are the child-0 columns shifted by two fine points.
// rp_shift = 0: fine point c sits at parent coordinate c/2;
// rp_shift = 1: half-cell offset (cell-centred), no coincident points.
std::vector<double> lagrange_matrix(int n, int cap, int rp)
{
    std::vector<double> A(static_cast<size_t>(n) * 2 * n, 0.0);
    auto at = [&](int p, int c) -> double& { return A[static_cast<size_t>(p) * 2 * n + c]; };
    for (int c = 1; c < 2 * n - 1; ++c)
    {
        if (rp == 0)
        {
            const int c_p = c - n;
            if (c % 2 == 0)
            {
                at(c / 2, c - 1) = 1.0;
                continue;
            }
            int bd_l, bd_r;
            if (c_p < -1)
            {
                bd_l = std::max(0, (c + 1) / 2 - cap / 2);
                bd_r = std::min(n, bd_l + cap);
            }
            else
            {
                bd_r = std::min(n, (c + 1) / 2 + cap / 2);
                bd_l = std::max(0, bd_r - cap);
            }
            long double sum = 0.0;
            for (int p = bd_l; p < bd_r; ++p)
            {
                long double m = 1.0;
                const int   p_p = p * 2 - n + 1;
                for (int l = bd_l; l < bd_r; ++l)
                    if (l != p) m /= -(long double)(2 * l - n + 1 - p_p);
                m /= (long double)(c_p - p_p + 1);
                at(p, c - 1) = (double)m;
                sum += m;
            }
            for (int p = 0; p < n; ++p) at(p, c - 1) /= (double)sum;
        }
        else
        {
            const double c_p = c - n + 0.5;
            int          bd_l, bd_r;
            if (c_p < -1)
            {
                bd_l = std::max(0, (c + 1) / 2 - cap / 2);
                bd_r = std::min(n, bd_l + cap);
            }
            else
            {
                bd_r = std::min(n, (c + 1) / 2 + cap / 2);
                bd_l = std::max(0, bd_r - cap);
            }
            long double sum = 0.0;
            for (int p = bd_l; p < bd_r; ++p)
            {
                long double m = 1.0;
                const int   p_p = p * 2 - n + 1;
                for (int l = bd_l; l < bd_r; ++l)
                    if (l != p) m /= -(long double)(2 * l - n + 1 - p_p);
                m /= (long double)(c_p - p_p);
                at(p, c - 1) = (double)m;
                sum += m;
            }
            for (int p = 0; p < n; ++p) at(p, c - 1) /= (double)sum;
        }
    }
    for (int c = 2 * n - 1; c > n - 1; --c)
        for (int p = 0; p < n; ++p) at(p, c) = at(p, c - 2);
    return A;
}
}  // namespace

std::vector<double> nli_matrix(int n, int pts_cap) { return lagrange_matrix(n, pts_cap, 0); }

std::vector<double> cc_interp_matrix(int n, int pts_cap, int relative_pos)
{
    return lagrange_matrix(n, pts_cap, relative_pos);
}

std::vector<double> cc_restrict_matrix(int n, int child_side, int relative_pos)
{
    // simple[0](i, 2i-1) = 1 ; simple[1](i, 2i-1) = simple[1](i, 2i) = 1/2, i in [1, n-2];
    // child 0 takes simple columns [1, n-2], child 1 columns [n-1, 2n-4].
    std::vector<double> S(static_cast<size_t>(n) * (2 * n - 2), 0.0);
    auto sat = [&](int i, int j) -> double& { return S[static_cast<size_t>(i) * (2 * n - 2) + j]; };
    for (int i = 1; i < n - 1; ++i)
    {
        if (relative_pos == 0) sat(i, 2 * i - 1) = 1.0;
        else
        {
            sat(i, 2 * i - 1) = 0.5;
            sat(i, 2 * i) = 0.5;
        }
    }
    std::vector<double> M(static_cast<size_t>(n) * n, 0.0);
    const int           off = child_side == 0 ? 1 : n - 1;
    for (int i = 0; i < n; ++i)
        for (int j = 1; j < n - 1; ++j) M[static_cast<size_t>(i) * n + j] = sat(i, off + j - 1);
    return M;
}

}  // namespace iblgf_amd
