// Synthetic vortex-ring workload (synth.cpp).
#pragma once

#include "octree.hpp"

#include <vector>

namespace iblgf_amd {

struct Ring
{
    double c[3];     // centre (physical units)
    double R;        // ring radius
    double a;        // Gaussian core radius
    double axis[3];  // unit axis
    double gamma;    // circulation
};

// Parameter block {nrings, source_max, refinement_factor, dx_base, then per
// ring: centre xyz, radius, core, axis xyz, circulation}.
constexpr int kRingParams = 9;

struct RingSet
{
    explicit RingSet(const double* params);
    // Refine where |omega| over the block grown by two cells exceeds
    // source_max * refinement_factor^diff (block_base in cells of its level).
    bool refine(const int base[3], int ext, int rl, int diff) const;
    // The components in comp_mask (bit c) of omega on the interior cells of
    // every leaf, out[c][slot][n^3] in slot order (level-major, key-sorted),
    // zeros elsewhere; threads over blocks.
    void vorticity(const Octree& t, int comp_mask, double* const out[3]) const;

    std::vector<Ring> rings;
    double            source_max = 1.0, refinement_factor = 0.9, dx_base = 1.0;
};

}  // namespace iblgf_amd
