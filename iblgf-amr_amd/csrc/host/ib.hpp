// Immersed-boundary geometry on the block octree: regularised delta
// functions, block overlap tests, the is_ib / extended_ib octant flags and
// the sparse delta weights that spread point forces to face nodes and
// interpolate face fields back to the points.
//
// Reference:
//   delta functions yang3 / yang4 / roma  include/iblgf/domain/ib.hpp:686-738
//   tensor-product delta, face offsets    ib.hpp:103-111; operators.hpp:844-861, 996-1015
//   ib_block_overlap                      ib.hpp:638-678
//   sphere points (Fibonacci spiral)      ib.hpp:179-200
//   update_ib_flag                        include/iblgf/domain/decomposition/server.hpp:544-575
//   update_ib_rank_and_infl               include/iblgf/domain/decomposition/client.hpp:1264-1320
#pragma once

#include "octree.hpp"

#include <vector>

namespace iblgf_amd {

enum DeltaKind
{
    DELTA_YANG3 = 0,  // reference default ("delta_func" = yang3), support 2
    DELTA_YANG4 = 1,  // support 2.5
    DELTA_ROMA = 2,   // 3-point, support 1.5
};
double delta_1d(int kind, double x);
double delta_support(int kind);

struct IbGeometry
{
    std::vector<double> x;           // N*3 physical coordinates
    int                 ib_level = 0;  // refinement level of the points (the finest)
    double              dx_base = 1.0;
    double              ddf_radius = 4.5;
    double              safety_dis = 4.0;  // 5/(Re dx_ib) + 1 in the reference
    int                 delta = DELTA_YANG3;

    int size() const { return static_cast<int>(x.size() / 3); }
    // Block [bbase, bbase+e) at refinement level rl overlaps point i's
    // region of radius class r (0: the block itself, 1: +ddf_radius+5,
    // 2: +ddf_radius+safety_dis+10), in cells of the IB level.
    bool block_overlap(int i, const int bbase[3], int e, int rl, int radius_level) const;
    bool block_overlap_any(const int bbase[3], int e, int rl, int radius_level) const;
};

// Fibonacci-spiral sphere of radius R with n points (ib.hpp:179-200).
std::vector<double> sphere_points(int n, double R);

// is_ib / extended_ib of the octants with data at the finest level.
void update_ib_flags(Octree& t, const IbGeometry& g);

// Sparse delta weights on the finest-level leaf blocks (slot numbering of
// the device mirror: slot_of maps octant -> slot; blocks n^3, x fastest).
//   interp rows r = 3*i + c: sum over (block in influence order, node in
//   block order) of field_c[slot*n^3 + q] * w;
//   spread rows: one per (c, node) touched, sum over points in index order.
struct IbStencil
{
    int              npoints = 0;
    std::vector<int> owned;  // 1 if a finest leaf block strictly contains the point
    std::vector<int> interp_ptr, interp_node;
    std::vector<double> interp_w;
    std::vector<int> spread_ptr, spread_dst /* {comp, index} */, spread_col /* 3*i+c */;
    std::vector<double> spread_w;
};
IbStencil build_ib_stencil(const Octree& t, const std::vector<int>& slot_of, const IbGeometry& g);

}  // namespace iblgf_amd
