// Block octree of the LGF fast-multipole path: 64-bit Morton keys, the
// pointer-free octant store, AMR refinement with 2:1 balance and correction
// buffers, 27-slot neighbour lists, <=189-slot influence lists and the
// AMR2AMR source/target masks.  Host side; integer work only.
//
// Reference behaviour restated here (file:line in IBLGF-AMR):
//   Morton key layout ............ include/iblgf/domain/octree/key.hpp:57-75, 119-132, 254-293, 447-469
//   Bitmasks (19 levels, 5+2 bits)  include/iblgf/domain/octree/bitmasks.hpp:56-77, 127-173
//   neighbour / influence keys ... include/iblgf/domain/octree/key.hpp:214-250
//   refine / try_2to1 ............ include/iblgf/domain/octree/tree.hpp:329-453
//   list construction ............ include/iblgf/domain/octree/tree.hpp:559-603
//   DFS order .................... include/iblgf/domain/octree/iterators.hpp:271-287
//   init_refine / mark_correction  include/iblgf/domain/domain.hpp:617-735, 878-943
//   fmm mask index ............... include/iblgf/domain/octree/octant.hpp:94-131
//   AMR2AMR masks ................ include/iblgf/fmm/fmm.hpp:392-680
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace iblgf_amd {

// ---------------------------------------------------------------------------
// Morton key.  bit0 = 1 (flag), bits 2..6 = level, coordinate bits of a
// level-L key occupy bits [64-3L, 64), x/y/z interleaved (x lowest).
// Coordinates are taken modulo 2^L (bits above L are shifted out).
// ---------------------------------------------------------------------------
struct Morton
{
    static constexpr int max_level = 19;
    static constexpr int flag_bits = 2;
    static constexpr int level_bits = 5;
    static constexpr int shift_to_coord = flag_bits + level_bits;

    static uint64_t spread(uint64_t w) noexcept
    {
        w &= 0x00000000001fffffULL;
        w = (w | w << 32) & 0x001f00000000ffffULL;
        w = (w | w << 16) & 0x001f0000ff0000ffULL;
        w = (w | w << 8) & 0x010f00f00f00f00fULL;
        w = (w | w << 4) & 0x10c30c30c30c30c3ULL;
        w = (w | w << 2) & 0x1249249249249249ULL;
        return w;
    }
    static int compact(uint64_t w) noexcept
    {
        w &= 0x1249249249249249ULL;
        w = (w ^ (w >> 2)) & 0x30c30c30c30c30c3ULL;
        w = (w ^ (w >> 4)) & 0xf00f00f00f00f00fULL;
        w = (w ^ (w >> 8)) & 0x00ff0000ff0000ffULL;
        w = (w ^ (w >> 16)) & 0x00ff00000000ffffULL;
        w = (w ^ (w >> 32)) & 0x00000000001fffffULL;
        return static_cast<int>(w);
    }
    static uint64_t make(int x, int y, int z, int level) noexcept
    {
        uint64_t idx = spread(static_cast<uint64_t>(static_cast<int64_t>(x)));
        idx |= spread(static_cast<uint64_t>(static_cast<int64_t>(y))) << 1;
        idx |= spread(static_cast<uint64_t>(static_cast<int64_t>(z))) << 2;
        idx <<= shift_to_coord;
        idx <<= (max_level - level) * 3;
        idx |= static_cast<uint64_t>(level) << flag_bits;
        idx |= 1ULL;
        return idx;
    }
    static int level(uint64_t k) noexcept
    {
        return static_cast<int>((k & 0x7cULL) >> flag_bits);
    }
    static void coord(uint64_t k, int c[3]) noexcept
    {
        // A level-0 key has no coordinate bits; the shift below would reach
        // 64 bits there (undefined in C++), so level 0 is (0,0,0) by definition.
        const int L = level(k);
        for (int d = 0; d < 3; ++d)
        {
            const int sh = d + shift_to_coord + (max_level - L) * 3;
            c[d] = sh >= 64 ? 0 : compact(k >> sh);
        }
    }
    // Top 3*l bits set (coordinate bits of levels 1..l).
    static uint64_t coord_mask(int l) noexcept
    {
        if (l <= 0) return 0ULL;
        return ~0ULL << (64 - 3 * l);
    }
    static uint64_t parent(uint64_t k) noexcept
    {
        const int L = level(k);
        if (L == 0) return k;
        return ((coord_mask(L - 1) & k) |
                   (static_cast<uint64_t>(L - 1) << flag_bits)) +
               1ULL;
    }
    static uint64_t child(uint64_t k, int i) noexcept
    {
        const int L = level(k);
        if (L == max_level) return k;
        return ((coord_mask(L) & k) |
                   (static_cast<uint64_t>(i)
                       << ((max_level - (L + 1)) * 3 + shift_to_coord)) |
                   (static_cast<uint64_t>(L + 1) << flag_bits)) +
               1ULL;
    }
    static int child_number(uint64_t k, int l) noexcept
    {
        if (l == 0) return -1;
        const uint64_t m = 7ULL << (64 - 3 * l);
        return static_cast<int>((m & k) >> (shift_to_coord + 3 * (max_level - l)));
    }
    // One beyond the largest key of the level: all coordinate bits set,
    // flag bit carried into bit 1.
    static uint64_t end_key(int l) noexcept
    {
        const uint64_t min_l = (static_cast<uint64_t>(l) << 2) | 1ULL;
        const uint64_t cmask = ~0ULL << shift_to_coord;
        return (min_l | cmask) + 1ULL;
    }
    static bool is_end(uint64_t k) noexcept { return k >= end_key(level(k)); }

    static uint64_t neighbor(uint64_t k, int dx, int dy, int dz) noexcept
    {
        int c[3];
        coord(k, c);
        return make(c[0] + dx, c[1] + dy, c[2] + dz, level(k));
    }
    // 27 keys, x fastest, offsets -1..1; slot (dx+1)+3(dy+1)+9(dz+1).
    static void neighbor_keys(uint64_t k, uint64_t out[27]) noexcept
    {
        int c[3];
        coord(k, c);
        const int L = level(k);
        int n = 0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx)
                    out[n++] = make(c[0] + dx, c[1] + dy, c[2] + dz, L);
    }
    // Children of the parent's 27 neighbours that are not adjacent to k,
    // in (parent-neighbour slot, child id) order.  Returns the count (<=216).
    static int influence_keys(uint64_t k, uint64_t* out) noexcept
    {
        if (level(k) == 0) return 0;
        uint64_t pn[27];
        neighbor_keys(parent(k), pn);
        int c[3];
        coord(k, c);
        int n = 0;
        for (int s = 0; s < 27; ++s)
        {
            if (is_end(pn[s])) continue;
            for (int ci = 0; ci < 8; ++ci)
            {
                const uint64_t ck = child(pn[s], ci);
                int cc[3];
                coord(ck, cc);
                if (std::abs(cc[0] - c[0]) > 1 || std::abs(cc[1] - c[1]) > 1 ||
                    std::abs(cc[2] - c[2]) > 1)
                    out[n++] = ck;
            }
        }
        return n;
    }
    // Smallest L with 2^L > max(m).
    static int minimum_level(int m) noexcept
    {
        int L = 0;
        while (m >= (1 << L)) ++L;
        return L;
    }
};

// FMM mask index map: STREAM 0, IB2xIB 1, xIB2IB 2, Laplacian_BC 3,
// IB2AMR 4+r, AMR2AMR 15+2r+nl, RefFourierStream 30+r, RefFourierLGF 40+r.
enum MaskType
{
    MASK_STREAM = 0,
    MASK_AMR2AMR = 1,
    MASK_IB2xIB = 2,
    MASK_xIB2IB = 3,
    MASK_IB2AMR = 4,
    MASK_Laplacian_BC = 5,
    MASK_RefFourierStream = 6,
    MASK_RefFourierLGF = 7
};
int fmm_mask_idx_gen(int type, int refinement_level = 0, int non_leaf_as_source = 0);
constexpr int kFmmMaxIdx = 50;

struct Octant
{
    uint64_t key = 0;
    int      level = 0;
    int      coord[3] = {0, 0, 0};
    int      parent = -1;
    int      children[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    bool     has_data = false;
    bool     leaf = false;
    bool     correction = false;
    bool     leaf_boundary = false;
    bool     physical = false;
    bool     aim_deletion = false;
    bool     is_ib = false;
    bool     extended_ib = false;
    int      neighbor[27];
    int      influence[189];
    int      influence_num = 0;
    uint64_t mask_src = 0;  // bit i: source mask of fmm mask index i
    uint64_t mask_tgt = 0;  // bit i: target mask of fmm mask index i

    bool mask(int idx, int which) const noexcept
    {
        const uint64_t m = which == 0 ? mask_src : mask_tgt;
        return (m >> idx) & 1ULL;
    }
    void set_mask(int idx, int which, bool v) noexcept
    {
        uint64_t& m = which == 0 ? mask_src : mask_tgt;
        if (v) m |= (1ULL << idx);
        else
            m &= ~(1ULL << idx);
    }
};

struct DomainSpec
{
    int    bd_base[3] = {0, 0, 0};
    int    bd_extent[3] = {0, 0, 0};
    int    block_extent = 14;
    double dx_base = 1.0;
    // Base-mesh regions in base-level cells (each divisible by block_extent).
    std::vector<std::array<int, 6>> regions;  // base xyz, extent xyz
};

// Refinement predicate: (octant, diff_level) -> 0 keep, 1 refine if the
// 2:1 balance allows it (ref_cond_), 2 refine unconditionally (a block that
// overlaps an immersed-boundary point, domain.hpp:887-891).
using RefineCond = std::function<int(const struct Octree&, int, int)>;

struct Octree
{
    std::vector<Octant>                  oct;  // oct[0] is the root
    std::unordered_map<uint64_t, int>    index;
    int                                  base_level = 0;
    int                                  depth = 0;
    int                                  block_extent = 14;
    int                                  base_shift[3] = {0, 0, 0};  // bd_base / block_extent
    int                                  key_box_base[3] = {0, 0, 0};
    int                                  key_box_extent[3] = {0, 0, 0};
    double                               dx_base = 1.0;
    bool                                 use_correction_buffer = true;
    std::vector<std::vector<int>>        level_map;  // key-sorted octants with data per level

    // -- construction --------------------------------------------------
    void build_base_mesh(const DomainSpec& spec);
    int  insert(uint64_t k);
    int  find(uint64_t k) const noexcept
    {
        auto it = index.find(k);
        return it == index.end() ? -1 : it->second;
    }
    int  refine_child(int o, int i);
    void give_data(int o) { oct[o].has_data = true; }
    int  refinement_level(int o) const noexcept { return oct[o].level - base_level; }

    // -- AMR ---------------------------------------------------------------
    bool try_2to1(uint64_t k, std::unordered_map<uint64_t, bool>& checklist) const;
    void refine(int o, bool ratio_2to1 = true);
    std::vector<uint64_t> unfound_neighbors(int o, bool correction_as_neighbors) const;
    void init_refine(int nRef, int level_up_max, const RefineCond& cond);
    // Decomposition::adapt_decoposition (decomposition.hpp:182-470) on one
    // rank: level_change (+1 refine, -1 coarsen) by leaf key; returns
    // base_mesh_update (base-level buffer blocks were added).  Lists are
    // rebuilt; masks are not (build them next).
    bool adapt(const std::unordered_map<uint64_t, int>& level_change);
    void delete_oct(int o);
    void deletion_reset_2to1(int o, std::unordered_set<int>& checklist);
    void delete_all_children();
    void mark_correction();
    void mark_leaf_boundary();

    // -- maps and lists ------------------------------------------------------
    std::vector<int> dfs_snapshot() const;
    void construct_leaf_maps(bool from_existing_flag);
    void construct_level_maps();
    void construct_lists();
    void neighbor_list_build(int o);
    void influence_list_build(int o);

    // -- masks -------------------------------------------------------------
    void fmm_lgf_mask_build(bool subtract_non_leaf);
    void fmm_dry(int base, bool non_leaf_as_source, bool subtract_non_leaf);
    void fmm_dry_init_base_level_masks(int base, bool nl, int idx, bool subtract_non_leaf);
    void fmm_upward_pass_masks(int base, int idx);
    void fmm_clean_no_inf_masks(int base, int idx);
    void fmm_ib_masks();
    void fmm_stream_masks();
    void fmm_ib2amr_masks();

    // -- geometry ------------------------------------------------------------
    // Block base (cells at the octant's refinement level, clamped at 0).
    void block_base(int o, int b[3]) const noexcept
    {
        int rl = refinement_level(o);
        if (rl < 0) rl = 0;
        for (int d = 0; d < 3; ++d) b[d] = (oct[o].coord[d] + base_shift[d] * (1 << rl)) * block_extent;
    }
    bool key_inside_box(uint64_t k) const noexcept
    {
        int c[3];
        Morton::coord(k, c);
        for (int d = 0; d < 3; ++d)
            if (c[d] < key_box_base[d] || c[d] > key_box_base[d] + key_box_extent[d] - 1) return false;
        return true;
    }
};

}  // namespace iblgf_amd
