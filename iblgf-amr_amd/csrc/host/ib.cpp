#include "ib.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <unordered_map>

namespace iblgf_amd {

namespace {
constexpr double kPi = 3.14159265358979323846;
}

// The three regularised delta kernels of ib.hpp:686-738, as functions of r = |x|.
// yang4's last branch multiplies its arcsine by the integer quotient 1/16,
// which is zero; the branch is kept as the reference evaluates it.
double delta_1d(int kind, double x)
{
    const double r = std::fabs(x);
    const double r2 = r * r;
    switch (kind)
    {
        case DELTA_ROMA:
            if (r > 1.5) return 0.0;
            if (r <= 0.5) return (1.0 + std::sqrt(-3. * r2 + 1.0)) / 3.0;
            return (5.0 - 3.0 * r - std::sqrt(1.0 - 3.0 * (1 - r) * (1 - r))) / 6.0;
        case DELTA_YANG3:
        {
            if (r > 2) return 0.0;
            const double s3 = std::sqrt(3);
            if (r <= 1.0)
                return 17.0 / 48.0 + s3 * kPi / 108.0 + r / 4.0 - r2 / 4.0 +
                       (1.0 - 2.0 * r) / 16 * std::sqrt(-12.0 * r2 + 12.0 * r + 1.0) -
                       s3 / 12.0 * std::asin(s3 / 2.0 * (2.0 * r - 1.0));
            return 55.0 / 48.0 - s3 * kPi / 108.0 - 13.0 * r / 12.0 + r2 / 4.0 +
                   (2.0 * r - 3.0) / 48.0 * std::sqrt(-12.0 * r2 + 36.0 * r - 23.0) +
                   s3 / 36.0 * std::asin(s3 / 2.0 * (2 * r - 3.0));
        }
        case DELTA_YANG4:
        {
            if (r > 2.5) return 0.0;
            if (r <= 0.5) return 3.0 / 8 + kPi / 32.0 - r2 / 4;
            if (r <= 1.5)
                return 1.0 / 4 + (1.0 - r) / 8.0 * std::sqrt(-2.0 + 8 * r - 4 * r2) -
                       1.0 / 8 * std::asin(std::sqrt(2) * (r - 1));
            return 17.0 / 16 - kPi / 64.0 - 3.0 / 4 * r + r2 / 8 + (r - 2.0) / 16.0 * std::sqrt(-14.0 + 16 * r - 4 * r2) +
                   0.0 * std::asin(std::sqrt(2) * (r - 2));
        }
        default: throw std::runtime_error("unknown delta function");
    }
}

double delta_support(int kind)
{
    switch (kind)
    {
        case DELTA_ROMA: return 1.5;
        case DELTA_YANG3: return 2.0;
        case DELTA_YANG4: return 2.5;
        default: throw std::runtime_error("unknown delta function");
    }
}

// ib_block_overlap (ib.hpp:648-678): the block is scaled to the IB level,
// grown by the radius (the integer descriptor truncates the fractional
// radius) and tested against the point scaled to that level.
bool IbGeometry::block_overlap(int i, const int bbase[3], int e, int rl, int radius_level) const
{
    if (rl > ib_level) throw std::runtime_error("block finer than the immersed-boundary level");
    const int f = 1 << (ib_level - rl);
    double    added = 0.0;
    if (radius_level == 2) added += ddf_radius + safety_dis + 10.0;
    else if (radius_level == 1)
        added += ddf_radius + 5.0;
    const int grow = static_cast<int>(2 * added), shift = static_cast<int>(added);
    const double factor = std::pow(2, ib_level) / dx_base;
    for (int d = 0; d < 3; ++d)
    {
        const int    lo = bbase[d] * f - shift;
        const int    ext = e * f + grow;
        const int    hi = lo + ext - 1;
        const double p = x[3 * i + d] * factor;
        if ((hi + 1) < p || lo >= p) return false;
    }
    return true;
}

bool IbGeometry::block_overlap_any(const int bbase[3], int e, int rl, int radius_level) const
{
    for (int i = 0; i < size(); ++i)
        if (block_overlap(i, bbase, e, rl, radius_level)) return true;
    return false;
}

std::vector<double> sphere_points(int n, double R)
{
    std::vector<double> p(3 * static_cast<size_t>(n));
    const double        lambda = (1. + std::sqrt(5)) / 2.0;
    for (int i = 0; i < n; ++i)
    {
        const double x = i + 0.5;
        const double phi = std::acos(1.0 - 2 * x / n);
        const double theta = 2 * kPi * x / lambda;
        p[3 * i + 0] = R * std::cos(theta) * std::sin(phi);
        p[3 * i + 1] = R * std::sin(theta) * std::sin(phi);
        p[3 * i + 2] = R * std::cos(phi);
    }
    return p;
}

void update_ib_flags(Octree& t, const IbGeometry& g)
{
    for (auto& o : t.oct)
    {
        o.is_ib = false;
        o.extended_ib = false;
    }
    const int L = t.depth - 1;
    if (L < 0 || L >= (int)t.level_map.size()) return;
    for (int i = 0; i < g.size(); ++i)
        for (int o : t.level_map[L])
        {
            if (!t.oct[o].has_data) continue;
            int b[3];
            t.block_base(o, b);
            const int rl = t.refinement_level(o);
            if (g.block_overlap(i, b, t.block_extent, rl, 2))
            {
                t.oct[o].extended_ib = true;
                if (t.oct[o].leaf && g.block_overlap(i, b, t.block_extent, rl, 1)) t.oct[o].is_ib = true;
            }
        }
}

IbStencil build_ib_stencil(const Octree& t, const std::vector<int>& slot_of, const IbGeometry& g)
{
    IbStencil st;
    const int N = g.size();
    st.npoints = N;
    st.owned.assign(N, 0);
    const int    L = t.depth - 1;
    const int    e = t.block_extent, n = e + 2;
    const long   n3 = (long)n * n * n;
    const double S = delta_support(g.delta);
    std::vector<int> blocks;  // finest leaf blocks with data, level order
    if (L >= 0 && L < (int)t.level_map.size())
        for (int o : t.level_map[L])
            if (t.oct[o].has_data && t.oct[o].leaf) blocks.push_back(o);
    const long max_index = (long)(slot_of.empty() ? 0 : *std::max_element(slot_of.begin(), slot_of.end()) + 1) * n3;
    if (max_index > 0x7fffffffL) throw std::runtime_error("field too large for 32-bit node indices");

    std::unordered_map<long, int>     row_of;  // comp*2^40 + index -> spread row
    std::vector<std::vector<int>>     srow_col;
    std::vector<std::vector<double>>  srow_w;
    st.interp_ptr.assign(1, 0);
    for (int i = 0; i < N; ++i)
    {
        std::vector<int>    infl;
        for (int o : blocks)
        {
            int b[3];
            t.block_base(o, b);
            const int rl = t.refinement_level(o);
            if (!g.block_overlap(i, b, e, rl, 1)) continue;
            infl.push_back(o);
            if (!st.owned[i] && g.block_overlap(i, b, e, rl, 0)) st.owned[i] = 1;
        }
        std::vector<int>    node[3];
        std::vector<double> wt[3];
        for (int o : infl)
        {
            int b[3];
            t.block_base(o, b);
            const int    rl = t.refinement_level(o);
            const double p = std::pow(2, rl);
            double       X[3];
            int          lo[3], hi[3];
            for (int d = 0; d < 3; ++d)
            {
                X[d] = g.x[3 * i + d] * p / g.dx_base;
                lo[d] = std::max(b[d], static_cast<int>(std::floor(X[d] - 0.5 - S)) - 1);
                hi[d] = std::min(b[d] + e - 1, static_cast<int>(std::ceil(X[d] + S)) + 1);
            }
            const int slot = slot_of[o];
            for (int kz = lo[2]; kz <= hi[2]; ++kz)
                for (int jy = lo[1]; jy <= hi[1]; ++jy)
                    for (int ix = lo[0]; ix <= hi[0]; ++ix)
                    {
                        const int    nc[3] = {ix, jy, kz};
                        double       dist[3];
                        bool         inside = true;
                        for (int d = 0; d < 3; ++d)
                        {
                            dist[d] = nc[d] - X[d];
                            if (std::fabs(dist[d]) >= g.ddf_radius + 1.0) inside = false;
                        }
                        if (!inside) continue;
                        const long q = (long)slot * n3 +
                                       ((long)(kz - b[2] + 1) * n + (jy - b[1] + 1)) * n + (ix - b[0] + 1);
                        for (int c = 0; c < 3; ++c)
                        {
                            double a[3];
                            for (int d = 0; d < 3; ++d) a[d] = dist[d] + (d == c ? 0.0 : 0.5);
                            const double w = delta_1d(g.delta, a[0]) * delta_1d(g.delta, a[1]) * delta_1d(g.delta, a[2]);
                            if (w == 0.0) continue;
                            node[c].push_back(static_cast<int>(q));
                            wt[c].push_back(w);
                            const long key = ((long)c << 40) | q;
                            auto       it = row_of.find(key);
                            int        r;
                            if (it == row_of.end())
                            {
                                r = static_cast<int>(srow_col.size());
                                row_of.emplace(key, r);
                                srow_col.emplace_back();
                                srow_w.emplace_back();
                                st.spread_dst.push_back(c);
                                st.spread_dst.push_back(static_cast<int>(q));
                            }
                            else
                                r = it->second;
                            srow_col[r].push_back(3 * i + c);
                            srow_w[r].push_back(w);
                        }
                    }
        }
        for (int c = 0; c < 3; ++c)
        {
            st.interp_node.insert(st.interp_node.end(), node[c].begin(), node[c].end());
            st.interp_w.insert(st.interp_w.end(), wt[c].begin(), wt[c].end());
            st.interp_ptr.push_back(static_cast<int>(st.interp_node.size()));
        }
    }
    st.spread_ptr.assign(1, 0);
    for (size_t r = 0; r < srow_col.size(); ++r)
    {
        st.spread_col.insert(st.spread_col.end(), srow_col[r].begin(), srow_col[r].end());
        st.spread_w.insert(st.spread_w.end(), srow_w[r].begin(), srow_w[r].end());
        st.spread_ptr.push_back(static_cast<int>(st.spread_col.size()));
    }
    return st;
}

}  // namespace iblgf_amd
