// Synthetic vortex-ring workload, evaluated natively: the refinement condition
// and the vorticity source of a set of Gaussian-core vortex rings.  They play
// the roles of the reference's test-setup source and refinement condition
// (tests/poisson/vortexrings.hpp:165-178, 471-520); the rings themselves are
// the seeded workload of the benchmark (thin rings with Gaussian cores).
#include "synth.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <thread>
#include <vector>

namespace iblgf_amd {

RingSet::RingSet(const double* p)
{
    if (!p) throw std::invalid_argument("ring parameters: null");
    const int nr = static_cast<int>(p[0]);
    if (nr < 0 || nr > 1 << 20 || p[0] != nr) throw std::invalid_argument("ring parameters: bad ring count");
    source_max = p[1];
    refinement_factor = p[2];
    dx_base = p[3];
    if (!(dx_base > 0)) throw std::invalid_argument("ring parameters: dx_base must be positive");
    rings.resize(nr);
    for (int r = 0; r < nr; ++r)
    {
        const double* q = p + 4 + kRingParams * r;
        Ring&         g = rings[r];
        for (int d = 0; d < 3; ++d)
        {
            g.c[d] = q[d];
            g.axis[d] = q[5 + d];
        }
        g.R = q[3];
        g.a = q[4];
        g.gamma = q[8];
        if (!(g.a > 0)) throw std::invalid_argument("ring parameters: core radius must be positive");
    }
}

// omega(P) of one ring: amplitude gamma / (pi a^2) exp(-rho^2 / a^2) along the
// circle's tangent axis x perp / |perp|, rho the distance to the core circle.
static inline void ring_add(const Ring& g, const double P[3], double w[3])
{
    double d[3], perp[3];
    for (int k = 0; k < 3; ++k) d[k] = P[k] - g.c[k];
    const double zax = d[0] * g.axis[0] + d[1] * g.axis[1] + d[2] * g.axis[2];
    for (int k = 0; k < 3; ++k) perp[k] = d[k] - zax * g.axis[k];
    const double rp = std::sqrt(perp[0] * perp[0] + perp[1] * perp[1] + perp[2] * perp[2]);
    if (!(rp > 0)) return;
    const double rho2 = (rp - g.R) * (rp - g.R) + zax * zax;
    const double amp = g.gamma / (M_PI * g.a * g.a) * std::exp(-rho2 / (g.a * g.a));
    const double et[3] = {g.axis[1] * perp[2] - g.axis[2] * perp[1], g.axis[2] * perp[0] - g.axis[0] * perp[2],
        g.axis[0] * perp[1] - g.axis[1] * perp[0]};
    for (int k = 0; k < 3; ++k) w[k] += amp * et[k] / rp;
}

// Distance from point x to the core circle of ring g.
static inline double circle_distance(const Ring& g, const double x[3])
{
    double d[3], perp[3];
    for (int k = 0; k < 3; ++k) d[k] = x[k] - g.c[k];
    const double zax = d[0] * g.axis[0] + d[1] * g.axis[1] + d[2] * g.axis[2];
    for (int k = 0; k < 3; ++k) perp[k] = d[k] - zax * g.axis[k];
    const double rp = std::sqrt(perp[0] * perp[0] + perp[1] * perp[1] + perp[2] * perp[2]);
    return std::hypot(rp - g.R, zax);
}

bool RingSet::refine(const int base[3], int ext, int rl, int diff) const
{
    const double dx = dx_base / std::ldexp(1.0, rl);
    const double thr = source_max * std::pow(refinement_factor, diff);
    // the block grown by two cells, as vortexrings.hpp:494
    double lo[3], hi[3], centre[3], hd = 0.0;
    for (int k = 0; k < 3; ++k)
    {
        lo[k] = (base[k] - 2 + 0.5) * dx;
        hi[k] = (base[k] + ext + 1 + 0.5) * dx;
        centre[k] = 0.5 * (lo[k] + hi[k]);
        hd += (hi[k] - lo[k]) * (hi[k] - lo[k]);
    }
    hd = 0.5 * std::sqrt(hd);
    // Rings whose core circle is farther from the grown block than the radius
    // where their peak falls below the threshold cannot refine it.
    std::vector<const Ring*> near;
    for (const Ring& g : rings)
    {
        const double amp = g.gamma / (M_PI * g.a * g.a);
        if (amp <= thr) continue;
        const double reach = g.a * std::sqrt(std::log(amp / thr));
        if (circle_distance(g, centre) - hd < reach) near.push_back(&g);
    }
    if (near.empty()) return false;
    const int m = ext + 4;
    for (int k = 0; k < m; ++k)
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i)
            {
                const double P[3] = {(base[0] - 2 + i + 0.5) * dx, (base[1] - 2 + j + 0.5) * dx,
                    (base[2] - 2 + k + 0.5) * dx};
                double w[3] = {0.0, 0.0, 0.0};
                for (const Ring& g : rings) ring_add(g, P, w);  // every ring, as the source
                if (std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]) > thr) return true;
            }
    return false;
}

void RingSet::vorticity(const Octree& t, int comp_mask, double* const out[3]) const
{
    const int              n = t.block_extent + 2, e = t.block_extent;
    const size_t           n3 = static_cast<size_t>(n) * n * n;
    std::vector<int>       slots;  // slot order: level-major, key-sorted
    for (const auto& lv : t.level_map)
        for (int o : lv) slots.push_back(o);
    const int nb = static_cast<int>(slots.size());
    for (int c = 0; c < 3; ++c)
        if (comp_mask & (1 << c)) std::fill(out[c], out[c] + n3 * nb, 0.0);
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto work = [&](unsigned tid) {
        for (int s = static_cast<int>(tid); s < nb; s += static_cast<int>(hw))
        {
            const int o = slots[s];
            if (!t.oct[o].leaf) continue;
            int b[3];
            t.block_base(o, b);
            const int    rl = std::max(t.refinement_level(o), 0);
            const double dx = dx_base / std::ldexp(1.0, rl);
            for (int k = 1; k <= e; ++k)
                for (int j = 1; j <= e; ++j)
                    for (int i = 1; i <= e; ++i)
                    {
                        const double P[3] = {(b[0] + i - 1 + 0.5) * dx, (b[1] + j - 1 + 0.5) * dx,
                            (b[2] + k - 1 + 0.5) * dx};
                        double w[3] = {0.0, 0.0, 0.0};
                        for (const Ring& g : rings) ring_add(g, P, w);
                        const size_t q = static_cast<size_t>(s) * n3 + (static_cast<size_t>(k) * n + j) * n + i;
                        for (int c = 0; c < 3; ++c)
                            if (comp_mask & (1 << c)) out[c][q] = w[c];
                    }
        }
    };
    std::vector<std::thread> pool;
    for (unsigned i = 1; i < hw; ++i) pool.emplace_back(work, i);
    work(0);
    for (auto& th : pool) th.join();
}

}  // namespace iblgf_amd
