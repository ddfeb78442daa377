// Host half of the C ABI of libiblgf_gpu.so: the octree, the Morton keys and
// the kernel tables (no device code; this file also builds on its own with a
// host compiler, e.g. under the sanitizers).  Declarations and reference
// citations are in include/iblgf_gpu.h.
#include "capi_common.hpp"

#include "host/chombo.hpp"
#include "host/synth.hpp"
#include "host/tables.hpp"

#include <algorithm>
#include <array>
#include <memory>
#include <mutex>
#include <vector>

using namespace iblgf_amd;

namespace iblgf_amd {
namespace capi {
namespace {
thread_local std::string g_last_error;
}  // namespace

int fail(int code, const std::string& msg, std::string* ctx_error)
{
    g_last_error = msg;
    if (ctx_error) *ctx_error = msg;
    return code;
}

const std::string& last_error() { return g_last_error; }

}  // namespace capi
}  // namespace iblgf_amd

namespace {
const LaplaceLgfTable& shared_table()
{
    static std::once_flag  once;
    static LaplaceLgfTable tab;
    std::call_once(once, [] { tab.build(100); });
    return tab;
}
}  // namespace

extern "C" {

const char* iblgf_last_error(void) { return capi::last_error().c_str(); }

int iblgf_tree_create(const int bd_base[3], const int bd_extent[3], int block_extent, double dx_base, int nregions,
    const int* regions, iblgf_tree** out)
{
    if (!bd_base || !bd_extent || !out || (nregions > 0 && !regions) || block_extent <= 0)
        return capi::fail(IBLGF_ERR_ARG, "iblgf_tree_create: invalid argument");
    return capi::guarded(nullptr, [&] {
        DomainSpec spec;
        for (int d = 0; d < 3; ++d)
        {
            spec.bd_base[d] = bd_base[d];
            spec.bd_extent[d] = bd_extent[d];
            if (bd_extent[d] % block_extent) throw std::invalid_argument("bounding box not divisible by block extent");
        }
        spec.block_extent = block_extent;
        spec.dx_base = dx_base;
        for (int r = 0; r < nregions; ++r)
        {
            std::array<int, 6> a;
            for (int k = 0; k < 6; ++k) a[k] = regions[6 * r + k];
            spec.regions.push_back(a);
        }
        auto t = std::make_unique<iblgf_tree>();
        t->tree.build_base_mesh(spec);
        t->tree.construct_level_maps();
        t->tree.construct_leaf_maps(false);
        t->tree.construct_lists();
        *out = t.release();
    });
}

int iblgf_tree_import(const int base_shift[3], int block_extent, int base_level, double dx_base, int nkeys,
    const uint64_t* keys, const uint32_t* flags, iblgf_tree** out)
{
    if (!base_shift || block_extent <= 0 || base_level < 0 || base_level > Morton::max_level || nkeys <= 0 ||
        !keys || !flags || !out)
        return capi::fail(IBLGF_ERR_ARG, "iblgf_tree_import: invalid argument");
    return capi::guarded(nullptr, [&] {
        auto   t = std::make_unique<iblgf_tree>();
        Octree& tr = t->tree;
        tr.block_extent = block_extent;
        tr.dx_base = dx_base;
        tr.base_level = base_level;
        tr.depth = base_level + 1;
        for (int d = 0; d < 3; ++d) tr.base_shift[d] = base_shift[d];
        tr.insert(Morton::make(0, 0, 0, 0));
        for (int i = 0; i < nkeys; ++i)
        {
            if ((keys[i] & 1ULL) == 0 || Morton::level(keys[i]) > Morton::max_level)
                throw std::invalid_argument("iblgf_tree_import: malformed key");
            const int o = tr.insert(keys[i]);
            if (tr.oct[o].has_data) throw std::invalid_argument("iblgf_tree_import: duplicate key");
            tr.oct[o].has_data = true;
            tr.oct[o].leaf = (flags[i] & 1u) != 0;
            tr.oct[o].correction = (flags[i] & 2u) != 0;
        }
        tr.construct_level_maps();
        tr.construct_lists();
        *out = t.release();
    });
}

int iblgf_tree_init_refine(iblgf_tree* t, int nref, int level_up_max, iblgf_refine_fn fn, void* user)
{
    if (!t || nref < 0 || level_up_max < 0 || (nref > 0 && !fn && !t->has_ib))
        return capi::fail(IBLGF_ERR_ARG, "iblgf_tree_init_refine: invalid argument");
    return capi::guarded(nullptr, [&] {
        RefineCond cond = [&](const Octree& tr, int o, int diff) {
            int b[3];
            tr.block_base(o, b);
            if (t->has_ib && t->ib.block_overlap_any(b, tr.block_extent, tr.refinement_level(o), 1)) return 2;
            const int r = fn ? fn(user, b, tr.block_extent, tr.refinement_level(o), diff) : 0;
            return r == 2 ? 2 : (r != 0 ? 1 : 0);
        };
        t->tree.init_refine(nref, level_up_max, cond);
        t->masks_built = false;
    });
}

int iblgf_tree_build_masks(iblgf_tree* t, int subtract_non_leaf)
{
    if (!t) return capi::fail(IBLGF_ERR_ARG, "iblgf_tree_build_masks: null tree");
    return capi::guarded(nullptr, [&] {
        if (t->has_ib) update_ib_flags(t->tree, t->ib);
        t->tree.fmm_lgf_mask_build(subtract_non_leaf != 0);
        t->tree.fmm_ib_masks();
        t->tree.fmm_stream_masks();
        t->tree.fmm_ib2amr_masks();
        t->masks_built = true;
        t->subtract_non_leaf = subtract_non_leaf != 0;
    });
}

int iblgf_tree_adapt(iblgf_tree* t, int n, const uint64_t* keys, const int* level_change, int* base_mesh_update)
{
    if (!t || n < 0 || (n > 0 && (!keys || !level_change)))
        return capi::fail(IBLGF_ERR_ARG, "iblgf_tree_adapt: invalid argument");
    return capi::guarded(nullptr, [&] {
        std::unordered_map<uint64_t, int> lc;
        for (int i = 0; i < n; ++i)
        {
            if (level_change[i] < -1 || level_change[i] > 1)
                throw std::invalid_argument("iblgf_tree_adapt: level changes are -1, 0 or +1");
            lc[keys[i]] = level_change[i];
        }
        const bool upd = t->tree.adapt(lc);
        if (base_mesh_update) *base_mesh_update = upd ? 1 : 0;
        t->masks_built = false;
    });
}

int iblgf_tree_set_ib(iblgf_tree* t, const double* points, int npoints, int ib_level, int delta, double safety_dis)
{
    if (!t || !points || npoints <= 0 || ib_level < 0 || delta < IBLGF_DELTA_YANG3 || delta > IBLGF_DELTA_ROMA)
        return capi::fail(IBLGF_ERR_ARG, "iblgf_tree_set_ib: invalid argument");
    return capi::guarded(nullptr, [&] {
        IbGeometry g;
        g.x.assign(points, points + 3 * static_cast<size_t>(npoints));
        g.ib_level = ib_level;
        g.dx_base = t->tree.dx_base;
        g.delta = delta;
        g.safety_dis = safety_dis;
        t->ib = std::move(g);
        t->has_ib = true;
        t->masks_built = false;
    });
}

int iblgf_tree_ib_stencil(const iblgf_tree* t, long* sizes4, int* interp_ptr, int* interp_node, double* interp_w,
    int* spread_ptr, int* spread_dst, int* spread_col, double* spread_w, int* owned)
{
    if (!t || !t->has_ib || !sizes4) return capi::fail(IBLGF_ERR_ARG, "iblgf_tree_ib_stencil: invalid argument");
    return capi::guarded(nullptr, [&] {
        std::vector<int> slot(t->tree.oct.size(), -1);
        int              k = 0;
        for (const auto& lm : t->tree.level_map)
            for (int o : lm) slot[o] = k++;
        const IbStencil st = build_ib_stencil(t->tree, slot, t->ib);
        sizes4[0] = static_cast<long>(st.interp_w.size());
        sizes4[1] = static_cast<long>(st.spread_ptr.size()) - 1;
        sizes4[2] = static_cast<long>(st.spread_w.size());
        sizes4[3] = st.npoints;
        auto put = [](auto* dst, const auto& v) {
            if (dst) std::copy(v.begin(), v.end(), dst);
        };
        put(interp_ptr, st.interp_ptr);
        put(interp_node, st.interp_node);
        put(interp_w, st.interp_w);
        put(spread_ptr, st.spread_ptr);
        put(spread_dst, st.spread_dst);
        put(spread_col, st.spread_col);
        put(spread_w, st.spread_w);
        put(owned, st.owned);
    });
}

int iblgf_tree_info(const iblgf_tree* t, int* nblocks, int* base_level, int* depth, int* block_extent,
    double* dx_base, int* nleaves)
{
    if (!t) return capi::fail(IBLGF_ERR_ARG, "iblgf_tree_info: null tree");
    int nb = 0, nl = 0;
    for (const auto& lm : t->tree.level_map)
        for (int o : lm)
        {
            ++nb;
            if (t->tree.oct[o].leaf) ++nl;
        }
    if (nblocks) *nblocks = nb;
    if (base_level) *base_level = t->tree.base_level;
    if (depth) *depth = t->tree.depth;
    if (block_extent) *block_extent = t->tree.block_extent;
    if (dx_base) *dx_base = t->tree.dx_base;
    if (nleaves) *nleaves = nl;
    return IBLGF_OK;
}

int iblgf_tree_export(const iblgf_tree* t, uint64_t* key, int* level, int* coord, int* bbase, uint32_t* flags,
    int* parent, int* children, int* neighbor, int* influence, int* influence_num, uint64_t* mask_src,
    uint64_t* mask_tgt)
{
    if (!t) return capi::fail(IBLGF_ERR_ARG, "iblgf_tree_export: null tree");
    return capi::guarded(nullptr, [&] {
        const Octree&    tr = t->tree;
        std::vector<int> slot(tr.oct.size(), -1), order;
        for (const auto& lm : tr.level_map)
            for (int o : lm)
            {
                slot[o] = static_cast<int>(order.size());
                order.push_back(o);
            }
        auto sl = [&](int o) { return o < 0 ? -1 : slot[o]; };
        for (size_t s = 0; s < order.size(); ++s)
        {
            const Octant& o = tr.oct[order[s]];
            if (key) key[s] = o.key;
            if (level) level[s] = o.level;
            int b[3];
            tr.block_base(order[s], b);
            for (int d = 0; d < 3; ++d)
            {
                if (coord) coord[3 * s + d] = o.coord[d];
                if (bbase) bbase[3 * s + d] = b[d];
            }
            if (flags)
                flags[s] = (o.leaf ? 1u : 0u) | (o.correction ? 2u : 0u) | (o.leaf_boundary ? 4u : 0u) |
                           (o.physical ? 8u : 0u) | (o.is_ib ? 16u : 0u) | (o.extended_ib ? 32u : 0u);
            if (parent) parent[s] = (o.parent >= 0 && tr.oct[o.parent].has_data) ? sl(o.parent) : -1;
            if (children)
                for (int c = 0; c < 8; ++c)
                    children[8 * s + c] = (o.children[c] >= 0 && tr.oct[o.children[c]].has_data) ? sl(o.children[c]) : -1;
            if (neighbor)
                for (int i = 0; i < 27; ++i) neighbor[27 * s + i] = sl(o.neighbor[i]);
            if (influence)
                for (int i = 0; i < 189; ++i) influence[189 * s + i] = i < o.influence_num ? sl(o.influence[i]) : -1;
            if (influence_num) influence_num[s] = o.influence_num;
            if (mask_src) mask_src[s] = o.mask_src;
            if (mask_tgt) mask_tgt[s] = o.mask_tgt;
        }
    });
}

void iblgf_tree_destroy(iblgf_tree* t) { delete t; }

// ---- output and restart files ------------------------------------------------

int iblgf_tree_write_restart(const iblgf_tree* t, const char* path)
{
    if (!t || !path) return capi::fail(IBLGF_ERR_ARG, "iblgf_tree_write_restart: invalid argument");
    return capi::guarded(nullptr, [&] { tree_write_restart(t->tree, path); });
}

int iblgf_tree_read_restart(const char* path, const int bd_base[3], const int bd_extent[3], int block_extent,
    double dx_base, iblgf_tree** out)
{
    if (!path || !bd_base || !bd_extent || block_extent <= 0 || !(dx_base > 0) || !out)
        return capi::fail(IBLGF_ERR_ARG, "iblgf_tree_read_restart: invalid argument");
    return capi::guarded(nullptr, [&] {
        std::vector<uint64_t> keys;
        std::vector<uint8_t>  leafs;
        tree_read_restart(path, keys, leafs);
        auto    t = std::make_unique<iblgf_tree>();
        Octree& tr = t->tree;
        tr.block_extent = block_extent;
        tr.dx_base = dx_base;
        int maxext = 0;
        for (int d = 0; d < 3; ++d)
        {
            if (bd_extent[d] % block_extent || bd_base[d] % block_extent)
                throw std::invalid_argument("iblgf_tree_read_restart: domain is not a whole number of blocks");
            tr.base_shift[d] = bd_base[d] / block_extent;
            maxext = std::max(maxext, bd_extent[d] / block_extent);
            tr.key_box_base[d] = 2;  // baseBlockBufferNumber
            tr.key_box_extent[d] = bd_extent[d] / block_extent - 4;
        }
        tr.base_level = Morton::minimum_level(maxext);
        tr.depth = tr.base_level + 1;
        tr.insert(Morton::make(0, 0, 0, 0));
        for (size_t i = 0; i < keys.size(); ++i)
        {
            if ((keys[i] & 1ULL) == 0 || Morton::level(keys[i]) > Morton::max_level)
                throw std::invalid_argument("iblgf_tree_read_restart: malformed key");
            const int o = tr.insert(keys[i]);
            if (tr.oct[o].has_data) throw std::invalid_argument("iblgf_tree_read_restart: duplicate key");
            tr.oct[o].has_data = true;
            tr.oct[o].leaf = leafs[i] != 0;
        }
        // Domain::restart_list_construct (domain.hpp:839-849).
        tr.construct_level_maps();
        tr.construct_lists();
        tr.construct_leaf_maps(true);
        tr.mark_correction();
        tr.mark_leaf_boundary();
        *out = t.release();
    });
}

int iblgf_chombo_write(const iblgf_tree* t, const char* path, int nfields, const char* const* names,
    const int* ncomp, const double* const* data, int include_correction, int base_correction_only, double time,
    double dt)
{
    if (!t || !path || nfields < 1 || !names || !ncomp || !data)
        return capi::fail(IBLGF_ERR_ARG, "iblgf_chombo_write: invalid argument");
    return capi::guarded(nullptr, [&] {
        std::vector<ChomboField> f;
        for (int i = 0; i < nfields; ++i)
        {
            if (!names[i] || ncomp[i] < 1 || !data[i]) throw std::invalid_argument("iblgf_chombo_write: bad field");
            f.push_back(ChomboField{names[i], ncomp[i], data[i]});
        }
        chombo_write(t->tree, path, f, include_correction != 0, base_correction_only != 0, time, dt);
    });
}

int iblgf_chombo_read(const iblgf_tree* t, const char* path, const char* name, int ncomp, int mesh_type,
    double* data)
{
    if (!t || !path || !name || ncomp < 1 || mesh_type < 0 || mesh_type > 3 || !data)
        return capi::fail(IBLGF_ERR_ARG, "iblgf_chombo_read: invalid argument");
    return capi::guarded(nullptr, [&] { chombo_read(t->tree, path, name, ncomp, mesh_type, data); });
}

uint64_t iblgf_key_make(int x, int y, int z, int level) { return Morton::make(x, y, z, level); }
int      iblgf_key_coord(uint64_t key, int* xyz)
{
    if (!xyz) return -1;
    Morton::coord(key, xyz);
    return Morton::level(key);
}
uint64_t iblgf_key_parent(uint64_t key) { return Morton::parent(key); }
uint64_t iblgf_key_child(uint64_t key, int i) { return Morton::child(key, i); }
int      iblgf_key_influence(uint64_t key, uint64_t* out) { return out ? Morton::influence_keys(key, out) : -1; }

int iblgf_lgf_table(int nmax, double* g)
{
    if (nmax < 0 || !g) return capi::fail(IBLGF_ERR_ARG, "iblgf_lgf_table: invalid argument");
    return capi::guarded(nullptr, [&] {
        LaplaceLgfTable tab;
        tab.build(nmax);
        std::copy(tab.g.begin(), tab.g.end(), g);
    });
}

double iblgf_lgf_get(int x, int y, int z) { return shared_table().get(x, y, z); }
double iblgf_if_kernel(int x, int y, int z, double alpha) { return if_kernel(x, y, z, alpha); }

int iblgf_lagrange_matrix(int n, int pts_cap, int relative_pos, double* out)
{
    if (n < 2 || pts_cap < 1 || !out || (relative_pos != 0 && relative_pos != 1))
        return capi::fail(IBLGF_ERR_ARG, "iblgf_lagrange_matrix: invalid argument");
    const auto A = relative_pos == 0 && pts_cap == 10 ? nli_matrix(n, 10) : cc_interp_matrix(n, pts_cap, relative_pos);
    std::copy(A.begin(), A.end(), out);
    return IBLGF_OK;
}

int iblgf_vortex_rings_refine(void* user, const int block_base[3], int block_extent, int refinement_level,
    int diff_level)
{
    // A refinement condition cannot report an error through its return value:
    // bad parameters refine nothing (iblgf_vortex_rings_vorticity rejects them).
    try
    {
        const RingSet rs(static_cast<const double*>(user));
        return rs.refine(block_base, block_extent, refinement_level, diff_level) ? 1 : 0;
    }
    catch (...)
    {
        return 0;
    }
}

int iblgf_vortex_rings_vorticity(const iblgf_tree* t, const double* params, int comp_mask, double* out_x,
    double* out_y, double* out_z)
{
    if (!t || !params || comp_mask <= 0 || comp_mask > 7 || ((comp_mask & 1) && !out_x) ||
        ((comp_mask & 2) && !out_y) || ((comp_mask & 4) && !out_z))
        return capi::fail(IBLGF_ERR_ARG, "iblgf_vortex_rings_vorticity: invalid argument");
    return capi::guarded(nullptr, [&] {
        const RingSet rs(params);
        double* const out[3] = {out_x, out_y, out_z};
        rs.vorticity(t->tree, comp_mask, out);
    });
}

}  // extern "C"
