// Host octree: construction, AMR refinement, lists and FMM masks.
// See octree.hpp for the reference file:line map.
#include "octree.hpp"

#include <algorithm>
#include <cstdlib>

namespace iblgf_amd {

int fmm_mask_idx_gen(int type, int refinement_level, int non_leaf_as_source)
{
    // include/iblgf/domain/octree/octant.hpp:94-131
    switch (type)
    {
        case MASK_STREAM: return 0;
        case MASK_IB2xIB: return 1;
        case MASK_xIB2IB: return 2;
        case MASK_Laplacian_BC: return 3;
        case MASK_IB2AMR: return refinement_level + 4;
        case MASK_AMR2AMR: return refinement_level * 2 + non_leaf_as_source + 15;
        case MASK_RefFourierStream: return refinement_level + 30;
        case MASK_RefFourierLGF: return refinement_level + 40;
        default: throw std::runtime_error("Wrong mask Idx");
    }
}

// ---------------------------------------------------------------------------
// Construction
// ---------------------------------------------------------------------------
int Octree::insert(uint64_t k)
{
    if (oct.empty())
    {
        Octant root;
        root.key = Morton::make(0, 0, 0, 0);
        root.level = 0;
        std::fill(root.neighbor, root.neighbor + 27, -1);
        std::fill(root.influence, root.influence + 189, -1);
        oct.push_back(root);
        index[root.key] = 0;
    }
    int       n = 0;
    const int L = Morton::level(k);
    for (int l = 1; l <= L; ++l)
    {
        const int cn = Morton::child_number(k, l);
        int       c = oct[n].children[cn];
        if (c < 0) c = refine_child(n, cn);
        n = c;
    }
    if (L + 1 > depth) depth = L + 1;
    return n;
}

int Octree::refine_child(int o, int i)
{
    if (oct[o].children[i] >= 0) return oct[o].children[i];
    Octant c;
    c.key = Morton::child(oct[o].key, i);
    c.level = oct[o].level + 1;
    Morton::coord(c.key, c.coord);
    c.parent = o;
    std::fill(c.neighbor, c.neighbor + 27, -1);
    std::fill(c.influence, c.influence + 189, -1);
    const int id = static_cast<int>(oct.size());
    oct.push_back(c);
    oct[o].children[i] = id;
    index[c.key] = id;
    return id;
}

void Octree::build_base_mesh(const DomainSpec& spec)
{
    // include/iblgf/domain/domain.hpp:166-211, 406-487
    block_extent = spec.block_extent;
    dx_base = spec.dx_base;
    const int e = spec.block_extent;
    std::vector<std::array<int, 3>> bases;
    for (const auto& r : spec.regions)
    {
        for (int d = 0; d < 3; ++d)
        {
            if (r[3 + d] % e) throw std::runtime_error("Domain: Extent of blocks are not evenly divisible");
            if (std::abs(r[d]) % e) throw std::runtime_error("Domain: Base of blocks are not evenly divisible");
        }
        for (int k = 0; k < r[5] / e; ++k)
            for (int j = 0; j < r[4] / e; ++j)
                for (int i = 0; i < r[3] / e; ++i)
                    bases.push_back({(r[0] + i * e) / e, (r[1] + j * e) / e, (r[2] + k * e) / e});
    }
    int maxext = 0;
    for (int d = 0; d < 3; ++d)
    {
        base_shift[d] = spec.bd_base[d] / e;
        maxext = std::max(maxext, spec.bd_extent[d] / e);
        key_box_base[d] = 2;  // baseBlockBufferNumber
        key_box_extent[d] = spec.bd_extent[d] / e - 4;
    }
    base_level = Morton::minimum_level(maxext);
    depth = base_level + 1;
    oct.clear();
    index.clear();
    insert(Morton::make(0, 0, 0, 0));
    for (auto& b : bases)
        insert(Morton::make(b[0] - base_shift[0], b[1] - base_shift[1], b[2] - base_shift[2], base_level));
    // The server instantiates a block on every octant of the initial tree.
    for (auto& o : oct) o.has_data = true;
}

// Pre-order DFS snapshot: root first, children in child-index order.
std::vector<int> Octree::dfs_snapshot() const
{
    std::vector<int> out;
    if (oct.empty()) return out;
    std::vector<int> st{0};
    while (!st.empty())
    {
        const int o = st.back();
        st.pop_back();
        out.push_back(o);
        for (int i = 7; i >= 0; --i)
            if (oct[o].children[i] >= 0) st.push_back(oct[o].children[i]);
    }
    return out;
}

// ---------------------------------------------------------------------------
// AMR: include/iblgf/domain/octree/tree.hpp:272-453
// ---------------------------------------------------------------------------
std::vector<uint64_t> Octree::unfound_neighbors(int o, bool correction_as_neighbors) const
{
    std::vector<uint64_t> keys;
    uint64_t              nk[27];
    Morton::neighbor_keys(oct[o].key, nk);
    for (int i = 0; i < 27; ++i)
    {
        if (Morton::is_end(nk[i])) continue;
        if (nk[i] == oct[o].key) continue;
        const int n = find(nk[i]);
        if (n < 0 || !oct[n].has_data) keys.push_back(nk[i]);
        else if (!correction_as_neighbors && oct[n].correction)
            keys.push_back(nk[i]);
    }
    return keys;
}

bool Octree::try_2to1(uint64_t k, std::unordered_map<uint64_t, bool>& checklist) const
{
    auto it = checklist.find(k);
    if (it != checklist.end()) return it->second;
    const int rf_l = Morton::level(k) - base_level;
    uint64_t  nk[27];
    Morton::neighbor_keys(k, nk);
    if (rf_l == 0)
    {
        for (int i = 0; i < 27; ++i)
        {
            if (nk[i] == k) continue;
            if (!key_inside_box(nk[i]))
            {
                checklist.emplace(k, false);
                return false;
            }
        }
        checklist.emplace(k, true);
        return true;
    }
    else if (rf_l > 0)
    {
        for (int i = 0; i < 27; ++i)
        {
            if (nk[i] == k) continue;
            if (!try_2to1(Morton::parent(nk[i]), checklist))
            {
                checklist.emplace(k, false);
                return false;
            }
        }
        checklist.emplace(k, true);
        return true;
    }
    return false;
}

void Octree::refine(int o, bool ratio_2to1)
{
    if (ratio_2to1)
    {
        for (int i = 0; i < 27; ++i)
        {
            const int n = oct[o].neighbor[i];
            if (n >= 0 && oct[n].has_data) oct[n].aim_deletion = false;
        }
        if (refinement_level(o) >= 0)
        {
            const auto keys = unfound_neighbors(o, false);
            for (uint64_t k : keys)
            {
                if (Morton::is_end(k)) continue;
                if (refinement_level(o) == 0)
                {
                    const int n = insert(k);
                    if (!oct[n].has_data) give_data(n);
                    oct[n].leaf = true;
                    oct[n].correction = false;
                    oct[n].leaf_boundary = false;
                    oct[n].aim_deletion = false;
                }
                else
                {
                    const uint64_t pk = Morton::parent(k);
                    int            pa = find(pk);
                    if (pa < 0)
                    {
                        pa = insert(pk);
                        give_data(pa);
                    }
                    refine(pa, ratio_2to1);
                }
            }
        }
    }
    for (int i = 0; i < 8; ++i)
    {
        const int c0 = oct[o].children[i];
        if (c0 >= 0 && oct[c0].has_data) continue;
        const int c = refine_child(o, i);
        if (!oct[c].has_data) give_data(c);
    }
    for (int i = 0; i < 8; ++i)
    {
        const int c = oct[o].children[i];
        oct[c].leaf = true;
        oct[c].correction = false;
        oct[c].leaf_boundary = false;
        oct[c].aim_deletion = false;
    }
    oct[o].leaf = false;
    oct[o].correction = false;
    oct[o].leaf_boundary = false;
    oct[o].aim_deletion = false;
    if (oct[o].level + 2 > depth) depth = oct[o].level + 2;
}

void Octree::init_refine(int nRef, int level_up_max, const RefineCond& cond)
{
    // include/iblgf/domain/domain.hpp:878-943
    construct_leaf_maps(false);
    std::unordered_map<uint64_t, bool> checklist;
    for (int l = 0; l < nRef; ++l)
    {
        // Live pre-order traversal: children of a node are pushed when the
        // node is popped, so blocks created under already-visited nodes are
        // not revisited (iterators.hpp:271-287).
        auto visit = [&](int cur) {
            if (refinement_level(cur) != l) return;
            if (!oct[cur].has_data) return;
            const int r = cond(*this, cur, nRef - l);
            if (r == 2) refine(cur, true);
            else if (r == 1 && try_2to1(oct[cur].key, checklist))
                refine(cur, true);
        };
        visit(0);  // the root is the first element, visited before it is popped
        std::vector<int> st{0};
        while (!st.empty())
        {
            const int cur = st.back();
            st.pop_back();
            for (int i = 7; i >= 0; --i)
                if (oct[cur].children[i] >= 0) st.push_back(oct[cur].children[i]);
            if (oct[cur].level == 0) continue;
            visit(cur);
        }
    }
    construct_level_maps();
    for (int g = 0; g < level_up_max; ++g)
    {
        for (int l = depth - 1; l >= base_level; --l)
        {
            const std::vector<int> lm = l < (int)level_map.size() ? level_map[l] : std::vector<int>();
            for (int o : lm)
                if (oct[o].leaf) refine(o, false);
        }
        construct_leaf_maps(true);
        construct_level_maps();
        construct_lists();
    }
    mark_correction();
    mark_leaf_boundary();
}

// Tree::delete_oct (tree.hpp:455-467): the block dropped, and the octant
// unlinked from its parent when it has no children.
void Octree::delete_oct(int o)
{
    oct[o].has_data = false;
    bool leaf_search = true;
    for (int i = 0; i < 8; ++i)
        if (oct[o].children[i] >= 0) leaf_search = false;
    if (!leaf_search) return;
    const int pa = oct[o].parent;
    if (pa < 0) return;
    for (int i = 0; i < 8; ++i)
        if (oct[pa].children[i] == o) oct[pa].children[i] = -1;
    index.erase(oct[o].key);
    oct[o].parent = -1;
}

// Tree::deletionReset_2to1 (tree.hpp:304-328)
void Octree::deletion_reset_2to1(int o, std::unordered_set<int>& checklist)
{
    if (checklist.count(o)) return;
    if (refinement_level(o) < 0) return;
    uint64_t nk[27];
    Morton::neighbor_keys(oct[o].key, nk);
    for (int i = 0; i < 27; ++i)
    {
        const int n = find(nk[i]);
        if (n < 0 || !oct[n].has_data) continue;
        oct[n].aim_deletion = false;
        if (oct[n].parent >= 0) deletion_reset_2to1(oct[n].parent, checklist);
    }
    checklist.insert(o);
}

// Domain::delete_all_children(deletion, true) (domain.hpp:512-568): a parent
// all of whose children with data are leaves or corrections aimed for
// deletion becomes the leaf, its children stay aimed for deletion; otherwise
// its children's deletion is cancelled.
void Octree::delete_all_children()
{
    for (int l = depth - 2; l >= base_level; --l)
    {
        if (l >= (int)level_map.size()) continue;
        for (int o : level_map[l])
        {
            bool leaf_search_data = true;
            for (int i = 0; i < 8; ++i)
            {
                const int c = oct[o].children[i];
                if (c >= 0 && oct[c].has_data) leaf_search_data = false;
            }
            if (leaf_search_data) continue;
            bool all = true;
            for (int i = 0; i < 8; ++i)
            {
                const int c = oct[o].children[i];
                if (c < 0 || !oct[c].has_data) continue;
                if (!oct[c].aim_deletion || (!oct[c].leaf && !oct[c].correction))
                {
                    all = false;
                    break;
                }
            }
            if (all)
            {
                for (int i = 0; i < 8; ++i)
                {
                    const int c = oct[o].children[i];
                    if (c < 0) continue;
                    oct[c].leaf = false;
                    oct[c].correction = false;
                    oct[c].leaf_boundary = false;
                    oct[c].aim_deletion = true;
                }
                oct[o].leaf = true;
                oct[o].correction = false;
                oct[o].leaf_boundary = false;
                oct[o].aim_deletion = false;
            }
            else
                for (int i = 0; i < 8; ++i)
                {
                    const int c = oct[o].children[i];
                    if (c >= 0) oct[c].aim_deletion = false;
                }
        }
    }
}

bool Octree::adapt(const std::unordered_map<uint64_t, int>& level_change)
{
    bool base_mesh_update = false;
    // corrections (not immersed-boundary blocks) are to be deleted
    for (int o : dfs_snapshot()) oct[o].aim_deletion = oct[o].correction && !oct[o].is_ib;
    std::unordered_map<int, int> change;
    for (const auto& kv : level_change)
    {
        const int o = find(kv.first);
        if (o < 0 || !oct[o].has_data) throw std::runtime_error("adapt: no block with key " + std::to_string(kv.first));
        change[o] = kv.second;
    }
    // leaves in key order (Tree::leaves_, a key-sorted map)
    std::vector<int> leaves;
    for (int o : dfs_snapshot())
        if (oct[o].has_data && oct[o].leaf) leaves.push_back(o);
    std::sort(leaves.begin(), leaves.end(), [&](int a, int b) { return oct[a].key < oct[b].key; });
    std::unordered_map<uint64_t, bool> checklist;
    std::vector<int>                   to_refine;
    for (int o : leaves)
    {
        auto it = change.find(o);
        const int lc = it == change.end() ? 0 : it->second;
        if (lc == 0) continue;
        if (lc < 0) oct[o].aim_deletion = true;
        else
        {
            if (!oct[o].is_ib && !try_2to1(oct[o].key, checklist)) continue;
            to_refine.push_back(o);
        }
    }
    for (int o : leaves)
        if (oct[o].is_ib || oct[o].extended_ib) oct[o].aim_deletion = false;
    // base-level blocks within baseBlockBufferNumber (2) of a kept block
    std::unordered_set<uint64_t> near;
    construct_level_maps();
    for (int o : level_map[base_level])
    {
        if (oct[o].aim_deletion || oct[o].correction) continue;
        int c[3];
        Morton::coord(oct[o].key, c);
        for (int dz = -2; dz <= 2; ++dz)
            for (int dy = -2; dy <= 2; ++dy)
                for (int dx = -2; dx <= 2; ++dx)
                {
                    const uint64_t nk = Morton::make(c[0] + dx, c[1] + dy, c[2] + dz, base_level);
                    if (!Morton::is_end(nk)) near.insert(nk);
                }
    }
    std::vector<uint64_t> near_sorted(near.begin(), near.end());
    std::sort(near_sorted.begin(), near_sorted.end());
    for (uint64_t nk : near_sorted)
    {
        int c[3];
        Morton::coord(nk, c);
        bool inside = true;
        for (int d = 0; d < 3; ++d)
            if (c[d] < key_box_base[d] - 2 || c[d] > key_box_base[d] + key_box_extent[d] + 1) inside = false;
        if (!inside) continue;
        int n = find(nk);
        if (n < 0 || !oct[n].has_data)
        {
            n = insert(nk);
            give_data(n);
            oct[n].leaf = true;
            oct[n].correction = false;
            oct[n].leaf_boundary = false;
            oct[n].aim_deletion = false;
            base_mesh_update = true;
        }
        oct[n].aim_deletion = false;
    }
    construct_level_maps();
    for (int o : to_refine)
        if (oct[o].leaf) refine(o, true);
    {
        std::unordered_set<int> reset;
        std::vector<int>        now;
        for (int o : dfs_snapshot())
            if (oct[o].has_data && oct[o].leaf) now.push_back(o);
        std::sort(now.begin(), now.end(), [&](int a, int b) { return oct[a].key < oct[b].key; });
        for (int o : now)
            if (refinement_level(o) > 0 && oct[o].parent >= 0) deletion_reset_2to1(oct[o].parent, reset);
    }
    construct_level_maps();
    // delete base-level leaves aimed for deletion
    for (int o : std::vector<int>(level_map[base_level]))
        if (oct[o].aim_deletion && oct[o].leaf) delete_oct(o);
    construct_level_maps();
    delete_all_children();
    construct_level_maps();
    construct_lists();
    mark_correction();
    mark_leaf_boundary();
    for (int o : dfs_snapshot())
        if (oct[o].has_data && oct[o].aim_deletion && refinement_level(o) > 0) delete_oct(o);
    construct_level_maps();
    for (int l = base_level - 1; l > 0; --l)
        for (int o : std::vector<int>(level_map[l]))
        {
            bool leaf_search = true;
            for (int i = 0; i < 8; ++i)
                if (oct[o].children[i] >= 0) leaf_search = false;
            if (leaf_search) delete_oct(o);
        }
    // every ancestor of a block holds a block
    for (int o : dfs_snapshot())
    {
        if (!oct[o].has_data) continue;
        for (int pa = oct[o].parent; pa >= 0 && !oct[pa].has_data; pa = oct[pa].parent) give_data(pa);
    }
    depth = 0;
    for (int o : dfs_snapshot())
        if (oct[o].has_data && oct[o].level + 1 > depth) depth = oct[o].level + 1;
    construct_level_maps();
    construct_lists();
    return base_mesh_update;
}

void Octree::mark_correction()
{
    // include/iblgf/domain/domain.hpp:617-735
    for (auto& o : oct) o.physical = false;
    for (int l = depth - 1; l >= base_level; --l)
    {
        if (l >= (int)level_map.size()) continue;
        for (int o : level_map[l])
        {
            if (!oct[o].has_data) continue;
            oct[o].physical = oct[o].leaf && !oct[o].aim_deletion;
            if (!oct[o].physical)
            {
                for (int i = 0; i < 8; ++i)
                {
                    const int c = oct[o].children[i];
                    if (c >= 0 && oct[c].has_data) oct[o].physical = oct[o].physical || oct[c].physical;
                }
            }
        }
    }
    if (use_correction_buffer)
    {
        const int d0 = depth;
        for (int l = base_level + 1; l < d0; ++l)
        {
            if (l >= (int)level_map.size()) continue;
            const std::vector<int> lm = level_map[l];
            for (int o : lm)
            {
                if (!oct[o].has_data) continue;
                if (!oct[o].physical) continue;
                const auto keys = unfound_neighbors(o, true);
                for (uint64_t k : keys)
                {
                    const int n = insert(k);
                    if (!oct[n].has_data) give_data(n);
                }
            }
        }
        construct_lists();
        construct_level_maps();
        construct_leaf_maps(true);
        for (int l = base_level + 1; l < depth; ++l)
        {
            if (l >= (int)level_map.size()) continue;
            for (int o : level_map[l])
            {
                if (!oct[o].physical) continue;
                for (int i = 0; i < 27; ++i)
                {
                    const int n = oct[o].neighbor[i];
                    if (n < 0 || !oct[n].has_data || oct[n].physical) continue;
                    oct[n].aim_deletion = false;
                    oct[n].correction = true;
                }
            }
        }
        construct_level_maps();
        construct_leaf_maps(true);
        construct_lists();
    }
    if (base_level < (int)level_map.size())
    {
        for (int o : level_map[base_level])
        {
            if (!oct[o].has_data) continue;
            if (!oct[o].physical) continue;
            bool all = true;
            for (int i = 0; i < 27; ++i)
            {
                const int n = oct[o].neighbor[i];
                if (n < 0 || !oct[n].has_data) all = false;
            }
            if (!all)
            {
                oct[o].correction = true;
                oct[o].aim_deletion = false;
                oct[o].leaf = true;
            }
            else
            {
                oct[o].correction = false;
                oct[o].aim_deletion = false;
            }
        }
    }
    construct_level_maps();
    construct_lists();
    construct_leaf_maps(true);
}

void Octree::mark_leaf_boundary()
{
    // include/iblgf/domain/domain.hpp:580-615
    for (auto& o : oct) o.leaf_boundary = false;
    const auto order = dfs_snapshot();
    for (int o : order)
    {
        if (!oct[o].has_data) continue;
        if (oct[o].leaf)
        {
            for (int i = 0; i < 27; ++i)
            {
                const int n = oct[o].neighbor[i];
                if (n < 0 || !oct[n].has_data || oct[n].correction || oct[n].leaf) continue;
                oct[o].leaf_boundary = true;
            }
        }
        else
        {
            if (oct[o].correction) continue;
            for (int i = 0; i < 27; ++i)
            {
                const int n = oct[o].neighbor[i];
                if (n < 0 || !oct[n].has_data || !oct[n].leaf) continue;
                oct[o].leaf_boundary = true;
            }
        }
        if (oct[o].leaf_boundary)
        {
            for (int i = 0; i < 8; ++i)
            {
                const int c = oct[o].children[i];
                if (c >= 0 && oct[c].has_data) oct[c].leaf_boundary = true;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Maps and lists: include/iblgf/domain/octree/tree.hpp:484-495, 559-603, 809-824
// ---------------------------------------------------------------------------
void Octree::construct_leaf_maps(bool from_existing_flag)
{
    if (from_existing_flag) return;  // leaf flags are the map
    for (int o : dfs_snapshot())
    {
        if (!oct[o].has_data) continue;
        bool leaf = true;
        for (int i = 0; i < 8; ++i)
            if (oct[o].children[i] >= 0) leaf = false;
        oct[o].leaf = leaf;
    }
}

void Octree::construct_level_maps()
{
    level_map.assign(Morton::max_level, {});
    for (int o = 0; o < (int)oct.size(); ++o)
        if (oct[o].has_data) level_map[oct[o].level].push_back(o);
    for (auto& lm : level_map)
        std::sort(lm.begin(), lm.end(), [&](int a, int b) { return oct[a].key < oct[b].key; });
}

void Octree::neighbor_list_build(int o)
{
    uint64_t nk[27];
    Morton::neighbor_keys(oct[o].key, nk);
    for (int i = 0; i < 27; ++i)
    {
        const int n = find(nk[i]);
        oct[o].neighbor[i] = (n < 0 || !oct[n].has_data) ? -1 : n;
    }
}

void Octree::influence_list_build(int o)
{
    std::fill(oct[o].influence, oct[o].influence + 189, -1);
    oct[o].influence_num = 0;
    if (oct[o].parent < 0) return;
    uint64_t  keys[216];
    const int nk = Morton::influence_keys(oct[o].key, keys);
    int       id = 0;
    for (int i = 0; i < nk; ++i)
    {
        const int n = find(keys[i]);
        if (n >= 0 && oct[n].has_data)
        {
            if (id >= 189) throw std::runtime_error("influence list overflow (>189)");
            oct[o].influence[id++] = n;
        }
    }
    oct[o].influence_num = id;
}

void Octree::construct_lists()
{
    for (int o : dfs_snapshot())
    {
        if (!oct[o].has_data) continue;
        neighbor_list_build(o);
        influence_list_build(o);
    }
}

// ---------------------------------------------------------------------------
// AMR2AMR masks: include/iblgf/fmm/fmm.hpp:392-680
// ---------------------------------------------------------------------------
void Octree::fmm_lgf_mask_build(bool subtract_non_leaf)
{
    for (int l = base_level; l < depth; ++l)
    {
        fmm_dry(l, false, subtract_non_leaf);
        fmm_dry(l, true, subtract_non_leaf);
    }
}

void Octree::fmm_dry(int base, bool nl, bool subtract_non_leaf)
{
    const int idx = fmm_mask_idx_gen(MASK_AMR2AMR, base - base_level, nl ? 1 : 0);
    fmm_dry_init_base_level_masks(base, nl, idx, subtract_non_leaf);
    fmm_upward_pass_masks(base, idx);
    fmm_clean_no_inf_masks(base, idx);
}

// Source (which = 0) and target (which = 1) masks of one FMM pass at its own
// level, for non_leaf_as_source = nl (rules of fmm.hpp:592-680).  A leaf or
// correction block is "leaf-like"; the subtract_non_leaf variants are built
// for export, the device solves refuse them.
void Octree::fmm_dry_init_base_level_masks(int base, bool nl, int idx, bool subtract_non_leaf)
{
    if (base >= (int)level_map.size()) return;
    for (int o : level_map[base])
    {
        Octant& it = oct[o];
        bool    src = false, tgt = false;
        if (it.has_data)
        {
            const bool leaf_like = it.leaf || it.correction;
            if (nl && subtract_non_leaf)
            {
                bool correction_child = false;
                for (int i = 0; i < 8; ++i)
                {
                    const int c = it.children[i];
                    if (c >= 0 && oct[c].has_data && oct[c].correction) correction_child = true;
                }
                src = !leaf_like;
                tgt = leaf_like ? correction_child : true;
            }
            else if (nl)
            {
                src = leaf_like;
                tgt = leaf_like ? it.leaf : true;
            }
            else if (subtract_non_leaf)
            {
                src = tgt = true;
            }
            else
            {
                src = !leaf_like;
                tgt = it.leaf;
            }
        }
        it.set_mask(idx, 0, src);
        it.set_mask(idx, 1, tgt);
    }
}

void Octree::fmm_upward_pass_masks(int base, int idx)
{
    for (int level = base - 1; level >= 0; --level)
    {
        if (level >= (int)level_map.size()) continue;
        for (int o : level_map[level])
        {
            oct[o].set_mask(idx, 0, false);
            for (int c = 0; c < 8; ++c)
            {
                const int ch = oct[o].children[c];
                if (ch >= 0 && oct[ch].has_data && oct[ch].mask(idx, 0))
                {
                    oct[o].set_mask(idx, 0, true);
                    break;
                }
            }
        }
        for (int o : level_map[level])
        {
            oct[o].set_mask(idx, 1, false);
            for (int c = 0; c < 8; ++c)
            {
                const int ch = oct[o].children[c];
                if (ch >= 0 && oct[ch].has_data && oct[ch].mask(idx, 1))
                {
                    oct[o].set_mask(idx, 1, true);
                    break;
                }
            }
        }
    }
}

void Octree::fmm_clean_no_inf_masks(int base, int idx)
{
    for (int level = 0; level < base; ++level)
    {
        if (level >= (int)level_map.size()) continue;
        for (int o : level_map[level])
        {
            Octant& it = oct[o];
            const int p = it.parent;
            if (it.mask(idx, 0))
            {
                // A source whose parent is a source is kept, and the target
                // check below is skipped for it as well (fmm.hpp:425-427).
                if (p >= 0 && oct[p].has_data && oct[p].mask(idx, 0)) continue;
                bool has_target = false;
                for (int i = 0; i < it.influence_num; ++i)
                {
                    const int n = it.influence[i];
                    if (n >= 0 && oct[n].has_data && oct[n].mask(idx, 1))
                    {
                        has_target = true;
                        break;
                    }
                }
                if (!has_target) it.set_mask(idx, 0, false);
            }
            if (it.mask(idx, 1))
            {
                if (p >= 0 && oct[p].has_data && oct[p].mask(idx, 1)) continue;
                bool has_source = false;
                for (int i = 0; i < it.influence_num; ++i)
                {
                    const int n = it.influence[i];
                    if (n >= 0 && oct[n].has_data && oct[n].mask(idx, 0))
                    {
                        has_source = true;
                        break;
                    }
                }
                if (!has_source) it.set_mask(idx, 1, false);
            }
        }
    }
}

void Octree::fmm_ib2amr_masks()
{
    // IB2AMR (FmmMaskBuilder::fmm_IB2AMR_mask, include/iblgf/fmm/fmm.hpp:146-195):
    // per refinement level, every block of the level is a target and the
    // blocks holding immersed-boundary points (or having a descendant that
    // does) are sources; parents are marked upward.  The reference marks the
    // ancestors is_ib in place; here the marking stays local to the masks.
    std::vector<char> ib(oct.size(), 0);
    for (size_t o = 0; o < oct.size(); ++o) ib[o] = oct[o].is_ib ? 1 : 0;
    for (int l = depth - 2; l >= base_level; --l)
    {
        if (l >= (int)level_map.size()) continue;
        for (int o : level_map[l])
            for (int c = 0; c < 8; ++c)
            {
                const int ch = oct[o].children[c];
                if (ch >= 0 && oct[ch].has_data && ib[ch]) ib[o] = 1;
            }
    }
    for (int l = depth - 1; l >= base_level; --l)
    {
        const int idx = fmm_mask_idx_gen(MASK_IB2AMR, l - base_level);
        if (idx >= fmm_mask_idx_gen(MASK_AMR2AMR, 0, 0)) throw std::runtime_error("too many levels for IB2AMR masks");
        for (auto& o : oct)
        {
            o.set_mask(idx, 0, false);
            o.set_mask(idx, 1, false);
        }
        if (l >= (int)level_map.size()) continue;
        for (int o : level_map[l])
        {
            oct[o].set_mask(idx, 1, true);
            if (ib[o]) oct[o].set_mask(idx, 0, true);
        }
        fmm_upward_pass_masks(l, idx);
    }
}

void Octree::fmm_stream_masks()
{
    // STREAM (FmmMaskBuilder::fmm_vortex_streamfun_mask, include/iblgf/fmm/fmm.hpp:199-233):
    // at the base level every block with data is a target and every
    // non-correction block a source; parents are marked upward.
    const int idx = fmm_mask_idx_gen(MASK_STREAM);
    for (auto& o : oct)
    {
        o.set_mask(idx, 0, false);
        o.set_mask(idx, 1, false);
    }
    if (base_level < (int)level_map.size())
        for (int o : level_map[base_level])
        {
            oct[o].set_mask(idx, 0, !oct[o].correction);
            oct[o].set_mask(idx, 1, true);
        }
    fmm_upward_pass_masks(base_level, idx);
}

void Octree::fmm_ib_masks()
{
    // IB2xIB / xIB2IB: include/iblgf/fmm/fmm.hpp:70-144.  Finest level only.
    const int l = depth - 1;
    for (int type : {MASK_IB2xIB, MASK_xIB2IB})
    {
        const int idx = fmm_mask_idx_gen(type);
        for (auto& o : oct)
        {
            o.set_mask(idx, 0, false);
            o.set_mask(idx, 1, false);
        }
        if (l < (int)level_map.size())
            for (int o : level_map[l])
            {
                if (oct[o].is_ib)
                {
                    oct[o].set_mask(idx, 0, true);
                    oct[o].set_mask(idx, 1, true);
                }
                if (oct[o].extended_ib)
                {
                    if (type == MASK_IB2xIB) oct[o].set_mask(idx, 1, true);
                    else
                        oct[o].set_mask(idx, 0, true);
                }
            }
        fmm_upward_pass_masks(l, idx);
        fmm_clean_no_inf_masks(l, idx);
    }
}

}  // namespace iblgf_amd
