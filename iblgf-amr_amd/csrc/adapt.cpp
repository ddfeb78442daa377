// Device mirror refresh after an adaptation of the host tree
// (Decomposition::adapt_decoposition, decomposition.hpp:182-560, and the
// field interpolation of Ifherk::adapt, ifherk.hpp:1018-1076).
//
// Kept blocks move from their old slot to their new one on the device (one
// block-copy launch per field, no host round trip); only the blocks of new
// leaves are computed, by interpolation from their parents; the plans and
// lists are rebuilt on the host and re-uploaded; the kernel spectra, which do
// not depend on the tree, stay resident.
#include "gpu_solver.hpp"
#include "kernels.hpp"
#include "ops.hpp"

#include <algorithm>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>

namespace iblgf_amd {

#define IBLGF_HIP_AD(x)                                                                               \
    do                                                                                                \
    {                                                                                                 \
        const hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_)); \
    } while (0)

void GpuSolver::coarsen(const int* f, int k, int mesh)
{
    if (mesh != kMeshCell && mesh != kMeshFace && mesh != kMeshEdge) throw std::invalid_argument("unknown mesh type");
    step_up(f, k, mesh);
    IBLGF_HIP_AD(hipGetLastError());
}

int GpuSolver::readapt(const Octree& t, const int* fields, const int* mesh, const int* comp, int nf)
{
    if (!uploaded_) throw std::runtime_error("tree not uploaded");
    if (ntrees_ != 1) throw std::invalid_argument("adaptation runs on a single tree");
    if (t.block_extent != e_ || t.base_level != base_level_ || t.dx_base != dx_base_)
        throw std::invalid_argument("the adapted tree must keep block extent, base level and dx_base");
    std::unordered_set<int> listed;
    for (int i = 0; i < nf; ++i)
    {
        if (field_is_f32(fields[i])) throw std::invalid_argument("adaptation carries FP64 fields only");
        if (mesh[i] != kMeshCell && mesh[i] != kMeshFace && mesh[i] != kMeshEdge)
            throw std::invalid_argument("unknown mesh type");
        if (comp[i] < 0 || comp[i] > 2) throw std::invalid_argument("field index out of range");
        if (!listed.insert(fields[i]).second) throw std::invalid_argument("a field is listed twice");
    }
    IBLGF_HIP_AD(hipSetDevice(device_));
    sync();
    // the old mirror: slot keys, leaf keys, field arrays
    const std::vector<uint64_t> old_key = key_;
    std::unordered_set<uint64_t> old_leaves;
    for (int s = 0; s < nb_; ++s)
        if (flags_[s] & 1u) old_leaves.insert(key_[s]);
    std::vector<DevArray<double>> old_fields = std::move(fields_);
    const std::vector<char>       old_f32 = field_f32_;
    // kernel spectra do not depend on the tree
    auto lap = std::move(ghat_lap_);
    auto lif = std::move(ghat_if_);
    auto lap32 = std::move(ghat_lap32_);
    auto lif32 = std::move(ghat_if32_);
    upload_tree(t);
    ghat_lap_ = std::move(lap);
    ghat_if_ = std::move(lif);
    ghat_lap32_ = std::move(lap32);
    ghat_if32_ = std::move(lif32);
    step_ready_ = false;  // the step's work fields were dropped with the old arrays
    // every field id again, at the new size
    fields_.clear();
    field_f32_.clear();
    for (size_t i = 0; i < old_fields.size(); ++i)
    {
        if (old_f32[i]) field_create_f32();
        else
            field_create();
    }
    // kept blocks
    std::unordered_map<uint64_t, int> new_slot;
    for (int s = 0; s < nb_; ++s) new_slot[key_[s]] = s;
    std::vector<int> pairs;
    for (int s = 0; s < (int)old_key.size(); ++s)
    {
        auto it = new_slot.find(old_key[s]);
        if (it != new_slot.end()) pairs.insert(pairs.end(), {s, it->second});
    }
    DevArray<int> d_pairs;
    d_pairs.upload(pairs);
    for (int i = 0; i < nf; ++i)
        launch::remap_blocks(stream_, n_, fields_[fields[i]].p, old_fields[fields[i]].p, d_pairs.p, d_pairs.size() / 2);
    // parents of new leaves (refinement level > 0), in new-leaf key order,
    // consecutive repeats dropped (they interpolate the same values again)
    std::vector<int> leaves;
    for (int s = 0; s < nb_; ++s)
        if ((flags_[s] & 1u) && level_[s] > base_level_ && !old_leaves.count(key_[s])) leaves.push_back(s);
    std::sort(leaves.begin(), leaves.end(), [&](int a, int b) { return key_[a] < key_[b]; });
    std::vector<int> parents;
    for (int s : leaves)
    {
        const int pa = parent_[s];
        if (pa >= 0 && (parents.empty() || parents.back() != pa)) parents.push_back(pa);
    }
    // one launch when no listed parent is a child of another listed parent
    // (the order does not matter then), else one launch per parent in order
    std::unordered_set<int> pset(parents.begin(), parents.end());
    bool independent = true;
    for (int pa : parents)
        if (parent_[pa] >= 0 && pset.count(parent_[pa])) independent = false;
    std::vector<std::vector<int>> batches;
    if (independent)
    {
        std::vector<int> all;
        for (int pa : std::unordered_set<int>(pset))
            for (int c = 0; c < 8; ++c)
                if (children_[8 * pa + c] >= 0) all.insert(all.end(), {children_[8 * pa + c], pa, c});
        if (!all.empty()) batches.push_back(std::move(all));
    }
    else
        for (int pa : parents)
        {
            std::vector<int> e;
            for (int c = 0; c < 8; ++c)
                if (children_[8 * pa + c] >= 0) e.insert(e.end(), {children_[8 * pa + c], pa, c});
            if (!e.empty()) batches.push_back(std::move(e));
        }
    std::vector<DevArray<int>> d_batches(batches.size());
    for (size_t b = 0; b < batches.size(); ++b) d_batches[b].upload(batches[b]);
    for (int i = 0; i < nf; ++i)
    {
        // Ifherk::adapt's loop over u's components: halo fills of every
        // component at every level, then this component's interpolation
        for (int l = depth_ - 2; l >= base_level_; --l)
            for (int j = 0; j < nf; ++j)
                launch::buffer_exchange(
                    stream_, n_, fields_[fields[j]].p, p_level_slots_[l].p, p_level_slots_[l].size(), d_neighbor_.p);
        Batch bt;
        bt.comp0 = comp[i];
        for (auto& d : d_batches)
            launch::transfer_scatter(stream_, n_, fields_[fields[i]].p, fields_[fields[i]].p, d.p, d.size() / 3,
                m_cc_interp_.p, false, bt, mesh[i]);
    }
    IBLGF_HIP_AD(hipGetLastError());
    sync();  // the old arrays are released on return
    return static_cast<int>(parents.size());
}

}  // namespace iblgf_amd
