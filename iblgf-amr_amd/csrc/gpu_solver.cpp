// Device mirror of the block octree and driver of the LGF solves.  See
// gpu_solver.hpp.
#include "gpu_solver.hpp"

#include "noalloc.hpp"

#include "ops.hpp"

#include "host/tables.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <unordered_map>

namespace iblgf_amd {

#define IBLGF_HIP(call)                                                                               \
    do                                                                                                \
    {                                                                                                 \
        hipError_t err_ = (call);                                                                     \
        if (err_ != hipSuccess)                                                                       \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(err_) + " at " #call); \
    } while (0)

template<class T>
void DevArray<T>::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
}
template<class T>
void DevArray<T>::alloc(size_t count)
{
    release();
    if (count == 0) return;
    IBLGF_HIP(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
    n = count;
}
template<class T>
void DevArray<T>::upload(const std::vector<T>& v)
{
    alloc(v.size());
    if (!v.empty()) IBLGF_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
}
template struct DevArray<int>;
template struct DevArray<double>;
template struct DevArray<double2>;
template struct DevArray<float2>;
template struct DevArray<float>;
template struct DevArray<uint16_t>;
template struct DevArray<MacGroup>;
template struct DevArray<MacSrc>;
template struct DevArray<uint4>;

// ---------------------------------------------------------------------------
// Kernel timer
// ---------------------------------------------------------------------------
KernelTimer::~KernelTimer()
{
    for (auto e : pool_) (void)hipEventDestroy(e);
}

void KernelTimer::enable(bool on)
{
    on_ = on;
    if (on && pool_.empty())
        for (int i = 0; i < 2048; ++i) pool_.push_back(next_event());
    used_ = 0;
    recs_.clear();
}

hipEvent_t KernelTimer::next_event()
{
    hipEvent_t e;
    IBLGF_HIP(hipEventCreate(&e));
    return e;
}

void KernelTimer::begin(hipStream_t s)
{
    if (!on_) return;
    if (used_ + 2 > pool_.size())
        for (int i = 0; i < 1024; ++i) pool_.push_back(next_event());
    IBLGF_HIP(hipEventRecord(pool_[used_], s));
}

void KernelTimer::end(hipStream_t s, int kind, long w)
{
    if (!on_) return;
    IBLGF_HIP(hipEventRecord(pool_[used_ + 1], s));
    recs_.push_back(Rec{kind, used_, w});
    used_ += 2;
}

void KernelTimer::collect()
{
    for (const auto& r : recs_)
    {
        float ms = 0.f;
        IBLGF_HIP(hipEventElapsedTime(&ms, pool_[r.ev], pool_[r.ev + 1]));
        total_ms[r.kind] += ms;
        launches[r.kind] += 1;
        work[r.kind] += r.work;
    }
    recs_.clear();
    used_ = 0;
}

void KernelTimer::reset()
{
    for (int k = 0; k < KC_COUNT; ++k)
    {
        total_ms[k] = 0.0;
        launches[k] = 0;
        work[k] = 0;
    }
    recs_.clear();
    used_ = 0;
}

GpuSolver::GpuSolver(int device) : device_(device)
{
    int count = 0;
    IBLGF_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) throw std::runtime_error("invalid device index");
    IBLGF_HIP(hipSetDevice(device));
    IBLGF_HIP(hipStreamCreate(&stream_));
    IBLGF_HIP(hipStreamCreate(&side_));
    IBLGF_HIP(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming));
    IBLGF_HIP(hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming));
}

GpuSolver::~GpuSolver()
{
    (void)hipSetDevice(device_);
    (void)hipStreamSynchronize(stream_);
    drop_graphs();
    passes_.clear();
    fields_.clear();
    (void)hipEventDestroy(ev_fork_);
    (void)hipEventDestroy(ev_join_);
    (void)hipStreamDestroy(side_);
    (void)hipStreamDestroy(stream_);
}

void GpuSolver::sync() { IBLGF_HIP(hipStreamSynchronize(stream_)); }

// ---------------------------------------------------------------------------
// Tree upload: slot table, plans, matrices, LGF table.
// ---------------------------------------------------------------------------
void GpuSolver::upload_tree(const Octree& t) { upload_forest({&t}); }

// Several independent trees in one slot table: slots are level-major, then by
// tree, then key-sorted, so every block list of a solve (and every launch)
// covers the blocks of all trees; lists and masks stay per tree.
void GpuSolver::upload_forest(const std::vector<const Octree*>& trees)
{
    if (trees.empty()) throw std::invalid_argument("no tree to upload");
    const Octree& t = *trees[0];
    for (const Octree* u : trees)
        if (!u || u->block_extent != t.block_extent || u->base_level != t.base_level || u->dx_base != t.dx_base)
            throw std::invalid_argument("forest trees must share block extent, base level and dx_base");
    IBLGF_HIP(hipSetDevice(device_));
    tree_ = t;
    ntrees_ = static_cast<int>(trees.size());
    n_ = t.block_extent + 2;
    if (!launch::supported_n(n_)) throw std::runtime_error("unsupported block extent for the device path");
    e_ = t.block_extent;
    n3_ = n_ * n_ * n_;
    const int P = 2 * n_ - 1, H = P / 2 + 1;
    nf_ = H * P * P;
    base_level_ = t.base_level;
    depth_ = 0;
    for (const Octree* u : trees) depth_ = std::max(depth_, u->depth);
    dx_base_ = t.dx_base;

    // slots: octants with data, level-major, then tree, key-sorted within
    oct_of_.clear();
    tree_of_.clear();
    lvl_begin_.assign(depth_ + 2, 0);
    for (int l = 0; l < depth_; ++l)
    {
        lvl_begin_[l] = static_cast<int>(oct_of_.size());
        for (int i = 0; i < ntrees_; ++i)
            if (l < (int)trees[i]->level_map.size())
                for (int o : trees[i]->level_map[l])
                {
                    oct_of_.push_back(o);
                    tree_of_.push_back(i);
                }
    }
    lvl_begin_[depth_] = lvl_begin_[depth_ + 1] = static_cast<int>(oct_of_.size());
    nb_ = static_cast<int>(oct_of_.size());
    forest_slot_of_.assign(ntrees_, {});
    for (int i = 0; i < ntrees_; ++i) forest_slot_of_[i].assign(trees[i]->oct.size(), -1);
    for (int s = 0; s < nb_; ++s) forest_slot_of_[tree_of_[s]][oct_of_[s]] = s;
    slot_of_ = forest_slot_of_[0];

    key_.resize(nb_);
    level_.resize(nb_);
    coord_.resize(3 * nb_);
    bbase_.resize(3 * nb_);
    flags_.resize(nb_);
    parent_.resize(nb_);
    children_.resize(8 * nb_);
    neighbor_.resize(27 * nb_);
    infl_.resize(189 * nb_);
    infl_num_.resize(nb_);
    msrc_.resize(nb_);
    mtgt_.resize(nb_);
    n_leaf_ = 0;
    for (int s = 0; s < nb_; ++s)
    {
        const Octree&           u = *trees[tree_of_[s]];
        const std::vector<int>& so = forest_slot_of_[tree_of_[s]];
        auto                    sl = [&](int o) { return o < 0 ? -1 : so[o]; };
        const Octant&           o = u.oct[oct_of_[s]];
        key_[s] = o.key;
        level_[s] = o.level;
        int b[3];
        u.block_base(oct_of_[s], b);
        for (int d = 0; d < 3; ++d)
        {
            coord_[3 * s + d] = o.coord[d];
            bbase_[3 * s + d] = b[d];
        }
        flags_[s] = (o.leaf ? 1u : 0u) | (o.correction ? 2u : 0u) | (o.is_ib ? 4u : 0u) | (o.extended_ib ? 8u : 0u);
        if (o.leaf) ++n_leaf_;
        parent_[s] = (o.parent >= 0 && u.oct[o.parent].has_data) ? sl(o.parent) : -1;
        for (int c = 0; c < 8; ++c)
        {
            const int ch = o.children[c];
            children_[8 * s + c] = (ch >= 0 && u.oct[ch].has_data) ? sl(ch) : -1;
        }
        for (int i = 0; i < 27; ++i) neighbor_[27 * s + i] = sl(o.neighbor[i]);
        infl_num_[s] = o.influence_num;
        for (int i = 0; i < 189; ++i) infl_[189 * s + i] = i < o.influence_num ? sl(o.influence[i]) : -1;
        msrc_[s] = o.mask_src;
        mtgt_[s] = o.mask_tgt;
    }
    d_neighbor_.upload(neighbor_);

    // transfer matrices: the FMM's Lagrange operators [2][n*n]; the
    // cell-centred restriction and interpolation [2 relative positions][2][n*n]
    {
        const auto A = nli_matrix(n_, 10);
        std::vector<double> an(2 * n_ * n_), in(2 * n_ * n_), cr(4 * n_ * n_), ci(4 * n_ * n_);
        for (int rel = 0; rel < 2; ++rel)
        {
            const auto C = cc_interp_matrix(n_, intrp_order_, rel);
            for (int b = 0; b < 2; ++b)
            {
                const auto   R = cc_restrict_matrix(n_, b, rel);
                const size_t o = static_cast<size_t>(rel * 2 + b) * n_ * n_;
                for (int p = 0; p < n_; ++p)
                    for (int j = 0; j < n_; ++j)
                    {
                        if (rel == 0)
                        {
                            const double a = A[(size_t)p * 2 * n_ + b * n_ + j];
                            an[b * n_ * n_ + p * n_ + j] = a;  // [parent][child]
                            in[b * n_ * n_ + j * n_ + p] = a;  // [child][parent]
                        }
                        ci[o + j * n_ + p] = C[(size_t)p * 2 * n_ + b * n_ + j];
                        cr[o + p * n_ + j] = R[(size_t)p * n_ + j];
                    }
            }
        }
        m_nli_anterp_.upload(an);
        m_nli_interp_.upload(in);
        m_cc_restrict_.upload(cr);
        m_cc_interp_.upload(ci);
    }

    // LGF table
    {
        LaplaceLgfTable tab;
        tab.build(lgf_nmax_);
        d_lgf_table_.upload(tab.g);
    }

    // fields and spectra
    fields_.clear();
    drop_graphs();
    work_ = Work{};
    reserve_components(1);

    // solver-level block lists
    std::vector<int> leaves, all(nb_), interior;
    for (int s = 0; s < nb_; ++s)
    {
        all[s] = s;
        if (flags_[s] & 1u) leaves.push_back(s);
    }
    for (int s = lvl_begin_[base_level_]; s < nb_; ++s)
        if (!(flags_[s] & 1u)) interior.push_back(s);
    p_leaves_.upload(leaves);
    p_all_.upload(all);
    p_nonleaf_.upload(interior);
    p_coarsify_.clear();
    p_level_slots_.clear();
    p_leaf_level_.clear();
    p_ccint_.clear();
    p_ccint_corr_.clear();
    p_kids_.clear();
    p_coarsify_.resize(depth_ + 1);
    p_level_slots_.resize(depth_ + 1);
    p_leaf_level_.resize(depth_ + 1);
    p_ccint_.resize(depth_ + 1);
    p_ccint_corr_.resize(depth_ + 1);
    p_kids_.resize(depth_ + 1);
    for (int l = 0; l < depth_; ++l)
    {
        std::vector<int> co, ls, ll, ci, cc, kids;
        for (int s = lvl_begin_[l]; s < lvl_begin_[l + 1]; ++s)
        {
            ls.push_back(s);
            if (flags_[s] & 1u) ll.push_back(s);
            bool any = false;
            std::vector<int> e(9, -1);
            e[0] = s;
            for (int c = 0; c < 8; ++c)
            {
                const int ch = children_[8 * s + c];
                if (ch < 0) continue;
                any = true;
                e[1 + c] = ch;
                ci.insert(ci.end(), {ch, s, c});
                if (flags_[ch] & 2u) cc.insert(cc.end(), {ch, s, c});
                kids.push_back(ch);
            }
            if (any) co.insert(co.end(), e.begin(), e.end());
        }
        p_coarsify_[l].upload(co);
        p_level_slots_[l].upload(ls);
        p_leaf_level_[l].upload(ll);
        p_ccint_[l].upload(ci);
        p_ccint_corr_[l].upload(cc);
        p_kids_[l].upload(kids);
    }
    passes_.clear();
    clean_lists_.clear();
    corr_level_.clear();
    smooth_lists_.clear();
    noncorr_lists_.clear();
    p_nonleaf_all_ = DevArray<int>{};
    p_leaf_ring_ = DevArray<int>{};
    d_ramp_ = DevArray<double>{};
    step_ready_ = false;
    ghat_lap_.clear();
    ghat_if_.clear();
    ghat_lap32_.clear();
    ghat_if32_.clear();
    uploaded_ = true;
}

int GpuSolver::field_create()
{
    DevArray<double> f;
    f.alloc(field_size());
    IBLGF_HIP(hipMemset(f.p, 0, f.n * sizeof(double)));
    fields_.push_back(std::move(f));
    field_f32_.push_back(0);
    return static_cast<int>(fields_.size()) - 1;
}

int GpuSolver::field_create_f32()
{
    DevArray<double> f;
    f.alloc((field_size() + 1) / 2);
    IBLGF_HIP(hipMemset(f.p, 0, f.n * sizeof(double)));
    fields_.push_back(std::move(f));
    field_f32_.push_back(1);
    return static_cast<int>(fields_.size()) - 1;
}

bool GpuSolver::field_is_f32(int id) const
{
    if (id < 0 || id >= (int)fields_.size()) throw std::runtime_error("invalid field id");
    return field_f32_[id] != 0;
}

double* GpuSolver::field_ptr(int id)
{
    if (field_is_f32(id)) throw std::invalid_argument("FP32 field: this operation takes FP64 fields");
    return fields_[id].p;
}

void* GpuSolver::field_data(int id)
{
    field_is_f32(id);  // id check
    return fields_[id].p;
}

void GpuSolver::field_upload_f32(int id, const float* host)
{
    if (!field_is_f32(id)) throw std::invalid_argument("FP64 field: upload FP64 values");
    IBLGF_HIP(hipMemcpyAsync(fields_[id].p, host, field_size() * sizeof(float), hipMemcpyHostToDevice, stream_));
    sync();
}

void GpuSolver::field_download_f32(int id, float* host)
{
    if (!field_is_f32(id)) throw std::invalid_argument("FP64 field: download FP64 values");
    IBLGF_HIP(hipMemcpyAsync(host, fields_[id].p, field_size() * sizeof(float), hipMemcpyDeviceToHost, stream_));
    sync();
}

void GpuSolver::field_upload(int id, const double* host)
{
    IBLGF_HIP(hipMemcpyAsync(field_ptr(id), host, field_size() * sizeof(double), hipMemcpyHostToDevice, stream_));
    sync();
}

void GpuSolver::field_download(int id, double* host)
{
    IBLGF_HIP(hipMemcpyAsync(host, field_ptr(id), field_size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    sync();
}

// Operator::maxnorm / blockRootMeanSquare (operators.hpp:1149-1167,
// 1351-1367) per block: the device reduces the sums of squares, the host takes
// the square roots (so the largest norm is the reference's bit for bit).
void GpuSolver::block_norms(const int* f, int k, double* maxnorm, double* rms)
{
    if (k < 1 || k > 3) throw std::invalid_argument("block norms of 1 to 3 components");
    const double* c[3] = {nullptr, nullptr, nullptr};
    for (int i = 0; i < k; ++i) c[i] = field_ptr(f[i]);
    DevArray<double> out;
    out.alloc(2 * static_cast<size_t>(nb_));
    launch::block_norms(stream_, n_, c, k, p_all_.p, nb_, out.p);
    std::vector<double> h(2 * static_cast<size_t>(nb_));
    IBLGF_HIP(hipMemcpyAsync(h.data(), out.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    sync();
    const double e3 = static_cast<double>(e_) * e_ * e_;
    for (int b = 0; b < nb_; ++b)
    {
        if (maxnorm) maxnorm[b] = std::sqrt(h[2 * b]);
        if (rms) rms[b] = std::sqrt(h[2 * b + 1] / e3);
    }
}

void GpuSolver::fill_zero(double* f, const Batch& b) { launch::zero_blocks(stream_, n_, f, p_all_.p, nb_, b); }

FieldTable GpuSolver::field_table(const int* ids, int k)
{
    static_assert(kMaxComponents <= kMaxFieldTable, "field tables hold every component");
    FieldTable t{};
    for (int c = 0; c < k; ++c)
    {
        t.p[c] = static_cast<double*>(field_data(ids[c]));
        if (field_f32_[ids[c]]) t.f32 |= 1ull << c;
    }
    return t;
}

FieldTable GpuSolver::work_table(double* base, int k) const
{
    FieldTable t{};
    for (int c = 0; c < k; ++c) t.p[c] = base + c * field_size();
    return t;
}

void GpuSolver::reserve_components(int k)
{
    if (k < 1 || k > kMaxComponents) throw std::invalid_argument("component count out of range");
    if (k <= work_.k) return;
    if (stream_) sync();  // solves in flight still use the old arrays
    drop_graphs();        // captured graphs point at the old arrays
    const size_t fs = field_size() * k, ss = static_cast<size_t>(nb_) * nf_ * k;
    Work w;
    for (DevArray<double>* a : {&w.src, &w.tgt, &w.fmm_s, &w.fmm_t})
    {
        a->alloc(fs);
        IBLGF_HIP(hipMemset(a->p, 0, fs * sizeof(double)));
    }
    w.sspec.alloc(ss);
    w.tspec.alloc(ss);
    IBLGF_HIP(hipMemset(w.sspec.p, 0, ss * sizeof(double2)));
    IBLGF_HIP(hipMemset(w.tspec.p, 0, ss * sizeof(double2)));
    w.k = k;
    work_ = std::move(w);
    wp_ = WorkPtrs{work_.src.p, work_.tgt.p, work_.fmm_s.p, work_.fmm_t.p, work_.sspec.p, work_.tspec.p};
}

Batch GpuSolver::batch(int k) const
{
    Batch b;
    b.k = k;
    b.mac_policy = mac_policy_;
    b.dft_grid = dft_grid_;
    b.fstride = static_cast<long>(field_size());
    b.sstride = static_cast<long>(nb_) * nf_;
    return b;
}

// ---------------------------------------------------------------------------
// Pass plans: Fmm::apply for base level L and non_leaf_as_source nl
// ---------------------------------------------------------------------------
void GpuSolver::build_pass(PassPlan& p, int L, int idx, bool nbonly)
{
    p.L = L;
    p.nl = (idx >= 15 && idx < 30) ? (idx - 15) % 2 : 0;
    p.neighbor_only = nbonly;
    p.idx = idx;
    auto      S = [&](int s) { return (msrc_[s] >> idx) & 1ULL; };
    auto      T = [&](int s) { return (mtgt_[s] >> idx) & 1ULL; };


    p.antrp.clear();
    p.antrp.resize(L + 1);
    p.intrp.clear();
    p.intrp.resize(L + 1);
    if (!nbonly)
    {
        for (int l = L - 1; l >= 0; --l)
        {
            std::vector<int> ent;
            for (int s = lvl_begin_[l]; s < lvl_begin_[l + 1]; ++s)
            {
                if (!S(s)) continue;
                std::vector<int> e(9, -1);
                e[0] = s;
                bool any = false;
                for (int c = 0; c < 8; ++c)
                {
                    const int ch = children_[8 * s + c];
                    if (ch < 0 || !S(ch)) continue;
                    e[1 + c] = ch;
                    any = true;
                    ++p.ntransfer;
                }
                if (any) ent.insert(ent.end(), e.begin(), e.end());
            }
            p.antrp[l].upload(ent);
        }
        p.ntransfer_antrp = p.ntransfer;
        // Level-L sources are read in place (interior, ring as zero) and the
        // coarser parents are written whole by the anterpolation; the coarser
        // sources it does not write (no source children) are cleared.
        std::vector<char> written(nb_, 0);
        for (int s = 0; s < lvl_begin_[L]; ++s)
            if (S(s))
                for (int c = 0; c < 8; ++c)
                {
                    const int ch = children_[8 * s + c];
                    if (ch >= 0 && S(ch)) written[s] = 1;
                }
        std::vector<int> zs;
        for (int l = L - 1; l >= 0; --l)
            for (int s = lvl_begin_[l]; s < lvl_begin_[l + 1]; ++s)
                if (S(s) && !written[s]) zs.push_back(s);
        p.zero_s.upload(zs);
        for (int l = 1; l < L; ++l)
        {
            std::vector<int> ent;
            for (int s = lvl_begin_[l]; s < lvl_begin_[l + 1]; ++s)
            {
                if (!T(s)) continue;
                for (int c = 0; c < 8; ++c)
                {
                    const int ch = children_[8 * s + c];
                    if (ch < 0 || !T(ch)) continue;
                    ent.insert(ent.end(), {ch, s, c});
                    ++p.ntransfer;
                }
            }
            p.intrp[l].upload(ent);
        }
    }

    // pairs per target and sibling groups per level
    std::vector<char> is_src(nb_, 0);
    std::vector<int>  inv;
    p.mac.clear();
    p.npairs = 0;
    const int lmin = nbonly ? L : 0;
    for (int lam = L; lam >= lmin; --lam)
    {
        struct GroupBuild
        {
            MacGroup                            g;
            std::vector<std::array<int, 3>>     ent;      // (slot, r, tmask)
            std::unordered_map<int, std::vector<int>> by_slot;
        };
        std::vector<GroupBuild>                gb;
        std::map<std::pair<int, uint64_t>, int> gidx;
        long                                   pairs_here = 0;
        for (int t = lvl_begin_[lam]; t < lvl_begin_[lam + 1]; ++t)
        {
            if (!T(t)) continue;
            std::vector<int> srcs;
            if (lam == L)
                for (int i = 0; i < 27; ++i)
                {
                    const int s = neighbor_[27 * t + i];
                    if (s >= 0 && S(s)) srcs.push_back(s);
                }
            if (!nbonly)
                for (int i = 0; i < infl_num_[t]; ++i)
                {
                    const int s = infl_[189 * t + i];
                    if (s >= 0 && S(s)) srcs.push_back(s);
                }
            if (srcs.empty()) continue;
            inv.push_back(t);
            // sibling groups are keyed by (tree, parent key): keys repeat across a forest
            const std::pair<int, uint64_t> pk(tree_of_[t], Morton::parent(key_[t]));
            auto                           it = gidx.find(pk);
            int            gi;
            if (it == gidx.end())
            {
                gi = static_cast<int>(gb.size());
                gidx[pk] = gi;
                GroupBuild b;
                std::fill(b.g.tslot, b.g.tslot + 8, -1);
                b.g.begin = b.g.count = b.g.nfull = 0;
                gb.push_back(std::move(b));
            }
            else
                gi = it->second;
            const int cx = coord_[3 * t] & 1, cy = coord_[3 * t + 1] & 1, cz = coord_[3 * t + 2] & 1;
            const int cb = cx | (cy << 1) | (cz << 2);
            gb[gi].g.tslot[cb] = t;
            for (int s : srcs)
            {
                int o[3], r[3];
                for (int d = 0; d < 3; ++d)
                {
                    o[d] = coord_[3 * s + d] - coord_[3 * t + d];
                    if (o[d] < -3 || o[d] > 3 || bbase_[3 * t + d] - bbase_[3 * s + d] != -o[d] * e_)
                        throw std::runtime_error(
                            "unsupported interaction: source outside the 7^3 block stencil (wrapped key)");
                }
                r[0] = o[0] + 2 + cx;
                r[1] = o[1] + 2 + cy;
                r[2] = o[2] + 2 + cz;
                for (int d = 0; d < 3; ++d)
                    if (r[d] < 0 || r[d] > 5) throw std::runtime_error("unsupported interaction: source outside group region");
                const int rr = r[0] + 6 * r[1] + 36 * r[2];
                // A pair listed twice (coarse-level Morton wrap-around makes a
                // parent neighbour alias another) is applied twice, as the
                // reference does: it goes into a further entry of the source.
                auto& ids = gb[gi].by_slot[s];
                bool  placed = false;
                for (int id : ids)
                {
                    auto& en = gb[gi].ent[id];
                    if (en[1] != rr) throw std::runtime_error("inconsistent source position in group");
                    if (!(en[2] & (1 << cb)))
                    {
                        en[2] |= (1 << cb);
                        placed = true;
                        break;
                    }
                }
                if (!placed)
                {
                    ids.push_back(static_cast<int>(gb[gi].ent.size()));
                    gb[gi].ent.push_back({s, rr, 1 << cb});
                }
                is_src[s] = 1;
                ++pairs_here;
            }
        }
        if (gb.empty()) continue;
        std::vector<MacGroup> groups;
        std::vector<MacSrc>   srcs;
        long                  m_nfull = 0, m_pairs_full = 0, m_pop[8] = {};
        for (auto& b : gb)
        {
            b.g.begin = static_cast<int>(srcs.size());
            std::vector<MacSrc> loc;
            for (auto& en : b.ent)
            {
                const int r = en[1];
                const int ob = (r % 6 + 1) + 7 * ((r / 6) % 6 + 1) + 49 * (r / 36 + 1);
                loc.push_back(MacSrc{en[0], ob | (en[2] << 9)});
            }
            // Entries acting on every target of the group first (the kernel
            // then computes all 8 children for them; absent children are
            // never stored), each part in source-position order.
            int gmask = 0;
            for (int c = 0; c < 8; ++c)
                if (b.g.tslot[c] >= 0) gmask |= 1 << c;
            std::stable_sort(loc.begin(), loc.end(), [gmask](const MacSrc& a, const MacSrc& c) {
                const bool fa = (a.meta >> 9) == gmask, fc = (c.meta >> 9) == gmask;
                if (fa != fc) return fa;
                return (a.meta & 511) < (c.meta & 511);
            });
            b.g.nfull = 0;
            for (const auto& en : loc) b.g.nfull += (en.meta >> 9) == gmask ? 1 : 0;
            for (const auto& en : loc)
            {
                const int pc = __builtin_popcount(static_cast<unsigned>(en.meta >> 9));
                if ((en.meta >> 9) == gmask)
                {
                    ++m_nfull;
                    m_pairs_full += pc;
                }
                else
                    ++m_pop[pc - 1];
            }
            srcs.insert(srcs.end(), loc.begin(), loc.end());
            b.g.count = static_cast<int>(loc.size());
            groups.push_back(b.g);
        }
        MacLevelPlan m;
        m.lambda = lam;
        m.d = L - lam;
        m.groups.upload(groups);
        m.srcs.upload(srcs);
        m.npairs = pairs_here;
        m.nentries = static_cast<long>(srcs.size());
        m.nfull = m_nfull;
        m.pairs_full = m_pairs_full;
        std::copy(m_pop, m_pop + 8, m.masked_pop);
        {
            std::vector<char> used(343, 0);
            for (const auto& e : srcs)
            {
                const int ob = e.meta & 511, tm = e.meta >> 9;
                for (int t = 0; t < 8; ++t)
                    if (tm & (1 << t)) used[ob - ((t & 1) + 7 * ((t >> 1) & 1) + 49 * ((t >> 2) & 1))] = 1;
            }
            std::vector<int> ol, row(343, 0);
            for (int o = 0; o < 343; ++o)
                if (used[o])
                {
                    row[o] = static_cast<int>(ol.size());
                    ol.push_back(o);
                }
            if (ol.empty()) ol.push_back(0);
            m.nused = static_cast<int>(ol.size());
            m.compact = m.nused <= kMacCompactMax;
            // LDS byte offsets of every entry's 8 kernel rows; children outside
            // the entry's mask read the zero row after the staged rows.
            const int        zrow = m.compact ? m.nused : 343;
            std::vector<uint4> ro(srcs.size());
            for (size_t i = 0; i < srcs.size(); ++i)
            {
                const int ob = srcs[i].meta & 511, tm = srcs[i].meta >> 9;
                unsigned  w[4] = {0, 0, 0, 0};
                for (int t = 0; t < 8; ++t)
                {
                    const int o = ob - ((t & 1) + 7 * ((t >> 1) & 1) + 49 * ((t >> 2) & 1));
                    const int r = (tm & (1 << t)) ? (m.compact ? row[o] : o) : zrow;
                    w[t >> 1] |= static_cast<unsigned>(r * kMacRowBytes) << (16 * (t & 1));
                }
                ro[i] = make_uint4(w[0], w[1], w[2], w[3]);
            }
            m.olist.upload(ol);
            m.rowoff.upload(ro);
            m.queue.upload(std::vector<int>(9, 0));
        }
        p.npairs += pairs_here;
        p.mac.push_back(std::move(m));
    }
    std::vector<int> fwd;
    for (int s = 0; s < nb_; ++s)
        if (is_src[s]) fwd.push_back(s);
    p.fwd.upload(fwd);
    p.nfwd_coarse = 0;
    while (p.nfwd_coarse < static_cast<int>(fwd.size()) && fwd[p.nfwd_coarse] < lvl_begin_[L]) ++p.nfwd_coarse;
    for (size_t i = p.nfwd_coarse; i < fwd.size(); ++i)
        if (fwd[i] >= lvl_begin_[L + 1]) throw std::logic_error("pass plan: sources finer than the pass level");
    p.inv.upload(inv);
    p.ninv_top = 0;
    while (p.ninv_top < static_cast<int>(inv.size()) && inv[p.ninv_top] >= lvl_begin_[L] &&
           inv[p.ninv_top] < lvl_begin_[L + 1])
        ++p.ninv_top;
    for (size_t i = p.ninv_top; i < inv.size(); ++i)
        if (inv[i] >= lvl_begin_[L]) throw std::logic_error("pass plan: level-L targets must lead the inverse list");
    // The level-L targets are added into the solve's target array directly;
    // the coarser targets are written whole by the inverse transform, and the
    // ones it does not write (no sources) are cleared for the interpolation
    // to read and add into.
    std::vector<char> written(nb_, 0);
    for (int t : inv) written[t] = 1;
    std::vector<int> zt;
    for (int l = L - 1; l >= 0; --l)
        for (int s = lvl_begin_[l]; s < lvl_begin_[l + 1]; ++s)
            if (T(s) && !written[s]) zt.push_back(s);
    p.zero_t.upload(zt);
}

PassPlan& GpuSolver::pass(int L, int nl, bool nbonly)
{
    return pass_idx(L, fmm_mask_idx_gen(MASK_AMR2AMR, L - base_level_, nl), nbonly);
}

PassPlan& GpuSolver::pass_idx(int L, int idx, bool nbonly)
{
    const int key = pass_key(L, idx, nbonly);
    auto      it = passes_.find(key);
    if (it != passes_.end()) return *it->second;
    auto p = std::make_unique<PassPlan>();
    build_pass(*p, L, idx, nbonly);
    auto& ref = *p;
    passes_[key] = std::move(p);
    return ref;
}

double2* GpuSolver::ghat_laplace(int d)
{
    auto it = ghat_lap_.find(d);
    if (it != ghat_lap_.end()) return it->second.p;
    const int  P = 2 * n_ - 1, H = P / 2 + 1;
    const long nsample = 343L * P * P * P, nleg = 343L * P * P * H;
    if (gen_sample_.n < (size_t)nsample) gen_sample_.alloc(nsample);
    if (gen_a_.n < (size_t)nleg) gen_a_.alloc(nleg);
    if (gen_b_.n < (size_t)nleg) gen_b_.alloc(nleg);
    DevArray<double2> g;
    g.alloc(static_cast<size_t>(343) * nf_);
    launch::lgf_spectra(stream_, n_, 0, d, d_lgf_table_.p, lgf_nmax_, nullptr, 0, gen_sample_.p, gen_a_.p,
        gen_b_.p, g.p);
    IBLGF_HIP(hipGetLastError());
    auto* ptr = g.p;
    ghat_lap_[d] = std::move(g);
    return ptr;
}

double2* GpuSolver::ghat_if(double alpha)
{
    auto it = ghat_if_.find(alpha);
    if (it != ghat_if_.end()) return it->second.p;
    const int  P = 2 * n_ - 1, H = P / 2 + 1;
    const long nsample = 343L * P * P * P, nleg = 343L * P * P * H;
    if (gen_sample_.n < (size_t)nsample) gen_sample_.alloc(nsample);
    if (gen_a_.n < (size_t)nleg) gen_a_.alloc(nleg);
    if (gen_b_.n < (size_t)nleg) gen_b_.alloc(nleg);
    const int           kmax = 3 * e_ + n_;
    std::vector<double> iv(kmax + 1);
    scaled_bessel_i(2.0 * alpha, kmax, iv.data());
    sync();  // the previous iv upload may still be read
    d_iv_.upload(iv);
    DevArray<double2> g;
    g.alloc(static_cast<size_t>(343) * nf_);
    launch::lgf_spectra(stream_, n_, 1, 0, nullptr, 0, d_iv_.p, kmax, gen_sample_.p, gen_a_.p, gen_b_.p, g.p);
    IBLGF_HIP(hipGetLastError());
    sync();
    auto* ptr = g.p;
    ghat_if_[alpha] = std::move(g);
    return ptr;
}

const void* GpuSolver::ghat(bool laplace, int d, double alpha)
{
    double2* g64 = laplace ? ghat_laplace(d) : ghat_if(alpha);
    if (precision_ == 0) return g64;
    DevArray<float2>* slot = nullptr;
    if (laplace)
    {
        auto it = ghat_lap32_.find(d);
        if (it != ghat_lap32_.end()) return it->second.p;
        slot = &ghat_lap32_[d];
    }
    else
    {
        auto it = ghat_if32_.find(alpha);
        if (it != ghat_if32_.end()) return it->second.p;
        slot = &ghat_if32_[alpha];
    }
    // generated in FP64, stored in FP32
    slot->alloc(static_cast<size_t>(343) * nf_);
    launch::to_float2(stream_, g64, slot->p, 343L * nf_);
    IBLGF_HIP(hipGetLastError());
    return slot->p;
}

void GpuSolver::set_precision(int p)
{
    if (p < kPrecFp64 || p > kPrecF16x3Forward)
        throw std::invalid_argument(
            "precision must be 0 (FP64), 1 (mixed FP32/FP64), 2-5 (matrix-core DFT legs) or 6 (matrix-core forward legs)");
    precision_ = p;
}

void GpuSolver::prepare_laplace()
{
    for (int l = base_level_; l < depth_; ++l)
        for (int nl = 0; nl < 2; ++nl)
        {
            PassPlan& p = pass(l, nl, false);
            for (auto& m : p.mac) ghat(true, m.d, 0.0);
        }
    sync();
}

void GpuSolver::prepare_if(double alpha_base)
{
    for (int l = base_level_; l < depth_; ++l)
    {
        for (int nl = 0; nl < 2; ++nl) pass(l, nl, true);
        ghat(false, 0, alpha_base * std::pow(4.0, l - base_level_));
    }
    sync();
}

// Fmm::apply (include/iblgf/fmm/fmm.hpp:729-870), for the b.k components held
// in the work arrays.
int GpuSolver::forward_legs() const { return precision_ == kPrecF16x3Forward ? kPrecF16x3 : precision_; }
int GpuSolver::inverse_legs() const { return precision_ == kPrecF16x3Forward ? kPrecMixed : precision_; }

void GpuSolver::run_pass(PassPlan& p, double scale, bool laplace, double alpha, const Batch& b)
{
    double*      fmm_s = wp_.fmm_s;
    double*      fmm_t = wp_.fmm_t;
    const int    L = p.L;
    const double P = 2.0 * n_ - 1.0, iscale = scale / (P * P * P);
    // Kernel spectra first: building one launches work on the solver stream.
    std::vector<const void*> gh(p.mac.size());
    for (size_t i = 0; i < p.mac.size(); ++i) gh[i] = ghat(laplace, p.mac[i].d, alpha);
    // Two chains over disjoint blocks.  Fine (solver stream): the level-L
    // sources' forward transform, read in place from the solve's source array
    // with their rings as zero (Fmm::apply copies interiors into zeroed
    // blocks), the level-L MAC and the inverse added into the level-L targets.
    // Coarse (side stream, when on): clears, anterpolation into fmm_s, the
    // coarser transforms and MACs, the inverse into fmm_t and the
    // interpolation down to level L-1.  They join before the last
    // interpolation, which adds into the level-L targets.  Issued in an order
    // that is also valid on one stream.
    const bool  coarse = !p.neighbor_only && L > 0;
    const bool  two = coarse && concurrent_;
    hipStream_t A = stream_, B = two ? side_ : stream_;
    if (two)
    {
        IBLGF_HIP(hipEventRecord(ev_fork_, A));
        IBLGF_HIP(hipStreamWaitEvent(B, ev_fork_, 0));
    }
    timer_.begin(B);
    launch::zero_blocks(B, n_, fmm_s, p.zero_s.p, p.zero_s.size(), b);
    launch::zero_blocks(B, n_, fmm_t, p.zero_t.p, p.zero_t.size(), b);
    timer_.end(B, KC_OTHER, 0);
    if (coarse)
    {
        timer_.begin(B);
        for (int l = L - 1; l >= 0; --l)
        {
            const bool top = l == L - 1;
            launch::transfer_gather(B, n_, top ? wp_.src : fmm_s, fmm_s, p.antrp[l].p, p.antrp[l].size() / 9,
                m_nli_anterp_.p, false, b, top);
        }
        timer_.end(B, KC_TRANSFER, p.ntransfer_antrp * b.k);
    }
    const int nfine = static_cast<int>(p.fwd.size()) - p.nfwd_coarse;
    timer_.begin(A);
    launch::forward_dft(A, n_, forward_legs(), fmm_s, p.fwd.p + p.nfwd_coarse, nfine, wp_.sspec, b, wp_.src, 0);
    timer_.end(A, KC_FORWARD, static_cast<long>(nfine) * b.k);
    timer_.begin(B);
    launch::forward_dft(B, n_, forward_legs(), fmm_s, p.fwd.p, p.nfwd_coarse, wp_.sspec, b);
    timer_.end(B, KC_FORWARD, static_cast<long>(p.nfwd_coarse) * b.k);
    for (size_t i = 0; i < p.mac.size(); ++i)
    {
        const MacLevelPlan& m = p.mac[i];
        hipStream_t         s = m.lambda == L ? A : B;
        timer_.begin(s);
        launch::spectral_mac(s, n_, precision_, m.groups.p, m.groups.size(), m.srcs.p, m.rowoff.p, m.olist.p,
            m.nused, m.compact, gh[i], wp_.sspec, wp_.tspec, b, m.queue.p);
        timer_.end(s, KC_MAC, m.npairs * b.k);
    }
    // Level-L targets (first in inv) add into the target array, coarser ones
    // are written to fmm_t and interpolated down.
    const int ntop = p.ninv_top, nrest = static_cast<int>(p.inv.size()) - ntop;
    timer_.begin(A);
    launch::inverse_dft(A, n_, inverse_legs(), wp_.tspec, p.inv.p, ntop, wp_.tgt, iscale, true, b);
    timer_.end(A, KC_INVERSE, static_cast<long>(ntop) * b.k);
    timer_.begin(B);
    launch::inverse_dft(B, n_, inverse_legs(), wp_.tspec, p.inv.p + ntop, nrest, fmm_t, iscale, false, b);
    timer_.end(B, KC_INVERSE, static_cast<long>(nrest) * b.k);
    if (coarse)
    {
        long w = 0;
        timer_.begin(B);
        for (int l = 1; l + 1 < L; ++l)
        {
            launch::transfer_scatter(B, n_, fmm_t, fmm_t, p.intrp[l].p, p.intrp[l].size() / 3, m_nli_interp_.p, true, b);
            w += static_cast<long>(p.intrp[l].size() / 3);
        }
        timer_.end(B, KC_TRANSFER, w * b.k);
        if (two)
        {
            IBLGF_HIP(hipEventRecord(ev_join_, B));
            IBLGF_HIP(hipStreamWaitEvent(A, ev_join_, 0));
        }
        if (L >= 2)
        {
            timer_.begin(A);
            launch::transfer_scatter(A, n_, fmm_t, wp_.tgt, p.intrp[L - 1].p, p.intrp[L - 1].size() / 3,
                m_nli_interp_.p, true, b);
            timer_.end(A, KC_TRANSFER, static_cast<long>(p.intrp[L - 1].size() / 3) * b.k);
        }
    }
}

void GpuSolver::set_concurrency(bool on)
{
    if (on == concurrent_) return;
    sync();
    drop_graphs();  // captured graphs hold the old launch order
    concurrent_ = on;
}

void GpuSolver::check_batch(const int* src, const int* dst, int k)
{
    if (!uploaded_) throw std::runtime_error("tree not uploaded");
    if (k < 1 || k > kMaxComponents) throw std::invalid_argument("component count out of range");
    for (int c = 0; c < k; ++c)
    {
        field_data(src[c]);
        field_data(dst[c]);
    }
    IBLGF_HIP(hipSetDevice(device_));
    reserve_components(k);
}

// PoissonSolver::apply_lgf (include/iblgf/solver/poisson/poisson.hpp:587-894)
void GpuSolver::apply_lgf(int src, int dst) { apply_lgf_batch(&src, &dst, 1); }

// Replays the batch from a captured graph when graphs are on (and kernel
// timing off): the first call with a key prepares the plans, kernel spectra
// and work arrays (host work and allocations stay outside the capture), then
// records the launches of `body` on the solver stream into a graph.
void GpuSolver::run_captured(const std::vector<int64_t>& key, const std::function<void()>& prepare,
    const std::function<void()>& body)
{
    if (!graphs_ || timer_.enabled())
    {
        body();
        return;
    }
    auto it = graph_cache_.find(key);
    if (it == graph_cache_.end())
    {
        prepare();
        sync();
        hipGraph_t g = nullptr;
        IBLGF_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        try
        {
            body();
        }
        catch (...)
        {
            hipGraph_t partial = nullptr;
            (void)hipStreamEndCapture(stream_, &partial);
            if (partial) (void)hipGraphDestroy(partial);
            throw;
        }
        IBLGF_HIP(hipStreamEndCapture(stream_, &g));
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        IBLGF_HIP(e);
        it = graph_cache_.emplace(key, exec).first;
    }
    IBLGF_HIP(hipGraphLaunch(it->second, stream_));
}

void GpuSolver::drop_graphs()
{
    for (auto& kv : graph_cache_) (void)hipGraphExecDestroy(kv.second);
    graph_cache_.clear();
}

void GpuSolver::set_graphs(bool on)
{
    if (!on) drop_graphs();
    graphs_ = on;
}

std::vector<int64_t> GpuSolver::batch_key(int kind, const int* src, const int* dst, int k, double alpha) const
{
    int64_t a = 0;
    static_assert(sizeof(a) == sizeof(alpha), "64-bit alpha");
    std::memcpy(&a, &alpha, sizeof(a));
    std::vector<int64_t> key = {kind, k, precision_, a};
    for (int c = 0; c < k; ++c)
    {
        key.push_back(src[c]);
        key.push_back(dst[c]);
    }
    return key;
}

static void check_mesh(int mesh)
{
    if (mesh != kMeshCell && mesh != kMeshFace && mesh != kMeshEdge) throw std::invalid_argument("unknown mesh type");
}

void GpuSolver::apply_lgf_batch(const int* src, const int* dst, int k, int mesh)
{
    check_mesh(mesh);
    check_batch(src, dst, k);
    auto key = batch_key(0, src, dst, k, 0.0);
    key.push_back(mesh);
    mesh_ = mesh;
    try
    {
        run_captured(key, [this] { prepare_laplace(); }, [this, src, dst, k] { lgf_batch_body(src, dst, k); });
    }
    catch (...)
    {
        mesh_ = kMeshCell;
        throw;
    }
    mesh_ = kMeshCell;
}

void GpuSolver::lgf_batch_body(const int* src, const int* dst, int k)
{
    const Batch b = batch(k);
    double* st = wp_.src;
    double* tt = wp_.tgt;
    timer_.begin(stream_);
    // Source: leaves copied whole, the other blocks at levels >= base cleared;
    // the coarsening then adds each parent's restricted children into it (a
    // leaf can have correction children, so parents are added into, not
    // overwritten).  Target: the base level is cleared for the passes to add
    // into; every finer block is first written whole by the interpolation from
    // its parent.
    launch::zero_blocks(stream_, n_, st, p_nonleaf_.p, p_nonleaf_.size(), b);
    launch::zero_blocks(stream_, n_, tt, p_level_slots_[base_level_].p, p_level_slots_[base_level_].size(), b);
    launch::copy_blocks_table(
        stream_, n_, work_table(st, k), field_table(src, k), p_leaves_.p, p_leaves_.size(), false, k);
    timer_.end(stream_, KC_OTHER, 0);
    timer_.begin(stream_);
    long ncoarse = 0;
    for (int ls = depth_ - 2; ls >= base_level_; --ls)
    {
        launch::transfer_gather(
            stream_, n_, st, st, p_coarsify_[ls].p, p_coarsify_[ls].size() / 9, m_cc_restrict_.p, true, b, false, mesh_);
        ncoarse += p_kids_[ls].size();
    }
    timer_.end(stream_, KC_TRANSFER, ncoarse * k);
    for (int l = base_level_; l < depth_; ++l)
    {
        const double dx = dx_base_ / std::pow(2.0, l - base_level_);
        run_pass(pass(l, 1, false), dx * dx, true, 0.0, b);
        timer_.begin(stream_);
        launch::transfer_scatter(
            stream_, n_, tt, tt, p_ccint_[l].p, p_ccint_[l].size() / 3, m_cc_interp_.p, false, b, mesh_);
        timer_.end(stream_, KC_TRANSFER, static_cast<long>(p_ccint_[l].size() / 3) * k);
        run_pass(pass(l, 0, false), dx * dx, true, 0.0, b);
        timer_.begin(stream_);
        launch::copy_blocks_table(stream_, n_, field_table(dst, k), work_table(tt, k), p_leaf_level_[l].p,
            p_leaf_level_[l].size(), false, k);
        if (l < depth_ - 1)
        {
            // Source correction of the children, added into the source in
            // place: the interpolated parent source (correction children),
            // then the stencil of the interpolated target (interior points).
            launch::buffer_exchange(
                stream_, n_, st, p_level_slots_[l].p, p_level_slots_[l].size(), d_neighbor_.p, b);
            launch::transfer_scatter(
                stream_, n_, st, st, p_ccint_corr_[l].p, p_ccint_corr_[l].size() / 3, m_cc_interp_.p, true, b, mesh_);
            const double dxc = dx / 2.0;
            launch::correction_stencil(
                stream_, n_, tt, st, p_kids_[l].p, p_kids_[l].size(), 1.0 / (dxc * dxc), 0.0, b);
        }
        timer_.end(stream_, KC_OTHER, 0);
    }
    IBLGF_HIP(hipGetLastError());
}

void GpuSolver::solve_sequence(const std::vector<SolveItem>& items)
{
    if (items.empty()) return;
    int kmax = 1;
    for (const SolveItem& it : items)
    {
        if (it.src.size() != it.dst.size() || it.src.empty()) throw std::invalid_argument("solve item fields");
        check_batch(it.src.data(), it.dst.data(), static_cast<int>(it.src.size()));
        kmax = std::max(kmax, static_cast<int>(it.src.size()));
    }
    reserve_components(kmax);
    std::vector<int64_t> key = {2, static_cast<int64_t>(items.size())};
    for (const SolveItem& it : items)
    {
        check_mesh(it.mesh);
        const auto k = batch_key(it.laplace ? 0 : 1, it.src.data(), it.dst.data(), static_cast<int>(it.src.size()),
            it.laplace ? 0.0 : it.alpha);
        key.insert(key.end(), k.begin(), k.end());
        key.push_back(it.mesh);
    }
    run_captured(
        key,
        [this, &items] {
            for (const SolveItem& it : items)
            {
                if (it.laplace) prepare_laplace();
                else
                    prepare_if(it.alpha);
            }
        },
        [this, &items] {
            for (const SolveItem& it : items)
            {
                const int k = static_cast<int>(it.src.size());
                mesh_ = it.mesh;
                if (it.laplace) lgf_batch_body(it.src.data(), it.dst.data(), k);
                else
                    if_batch_body(it.src.data(), it.dst.data(), k, it.alpha);
            }
            mesh_ = kMeshCell;
        });
    mesh_ = kMeshCell;
}

void GpuSolver::set_launch_shapes(int mac_policy, int dft_grid)
{
    if (mac_policy < 0 || mac_policy > 2 || dft_grid < 0 || dft_grid > 1)
        throw std::invalid_argument("launch shape out of range");
    if (mac_policy == mac_policy_ && dft_grid == dft_grid_) return;
    sync();
    drop_graphs();  // captured graphs hold the old launch shapes
    mac_policy_ = mac_policy;
    dft_grid_ = dft_grid;
}

std::string GpuSolver::autotune(const int* src, const int* dst, int k, bool laplace, double alpha_base, int reps)
{
    if (reps < 1) throw std::invalid_argument("autotune: reps must be positive");
    check_batch(src, dst, k);
    if (laplace) prepare_laplace();
    else
        prepare_if(alpha_base);
    const bool graphs = graphs_, timing = timer_.enabled();
    graphs_ = false;
    timer_.enable(false);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    IBLGF_HIP(hipEventCreate(&e0));
    IBLGF_HIP(hipEventCreate(&e1));
    auto solve = [&] {
        if (laplace) lgf_batch_body(src, dst, k);
        else
            if_batch_body(src, dst, k, alpha_base);
    };
    std::string report;
    double      best = -1.0;
    int         bm = mac_policy_, bd = dft_grid_;
    bool        bc = concurrent_;
    try
    {
        for (int conc = 1; conc >= 0; --conc)
            for (int mp = 0; mp < 3; ++mp)
                for (int dg = 0; dg < 2; ++dg)
                {
                    set_concurrency(conc != 0);
                    set_launch_shapes(mp, dg);
                    solve();  // warm-up
                    IBLGF_HIP(hipEventRecord(e0, stream_));
                    for (int r = 0; r < reps; ++r) solve();
                    IBLGF_HIP(hipEventRecord(e1, stream_));
                    IBLGF_HIP(hipEventSynchronize(e1));
                    float ms = 0.f;
                    IBLGF_HIP(hipEventElapsedTime(&ms, e0, e1));
                    const double per = ms / reps;
                    report += "mac_policy=" + std::to_string(mp) + " dft_grid=" + std::to_string(dg) +
                              " streams=" + std::to_string(conc ? 2 : 1) + " ms_per_call=" + std::to_string(per) + "\n";
                    if (best < 0 || per < best)
                    {
                        best = per;
                        bm = mp;
                        bd = dg;
                        bc = conc != 0;
                    }
                }
    }
    catch (...)
    {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        graphs_ = graphs;
        timer_.enable(timing);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    set_concurrency(bc);
    set_launch_shapes(bm, bd);
    graphs_ = graphs;
    timer_.enable(timing);
    report += "chosen: mac_policy=" + std::to_string(bm) + " dft_grid=" + std::to_string(bd) +
              " streams=" + std::to_string(bc ? 2 : 1) + " ms_per_call=" + std::to_string(best) + "\n";
    return report;
}

// PoissonSolver::apply_if (include/iblgf/solver/poisson/poisson.hpp:435-494)
void GpuSolver::apply_if(int src, int dst, double alpha_base) { apply_if_batch(&src, &dst, 1, alpha_base); }

void GpuSolver::apply_if_batch(const int* src, const int* dst, int k, double alpha_base, int mesh)
{
    check_mesh(mesh);
    check_batch(src, dst, k);
    auto key = batch_key(1, src, dst, k, alpha_base);
    key.push_back(mesh);
    mesh_ = mesh;
    try
    {
        run_captured(key, [this, alpha_base] { prepare_if(alpha_base); },
            [this, src, dst, k, alpha_base] { if_batch_body(src, dst, k, alpha_base); });
    }
    catch (...)
    {
        mesh_ = kMeshCell;
        throw;
    }
    mesh_ = kMeshCell;
}

bool GpuSolver::has_ghat(bool laplace, int d, double alpha) const
{
    if (precision_ == 0) return laplace ? ghat_lap_.count(d) > 0 : ghat_if_.count(alpha) > 0;
    return laplace ? ghat_lap32_.count(d) > 0 : ghat_if32_.count(alpha) > 0;
}

bool GpuSolver::prepared(bool laplace, double alpha_base) const
{
    if (!uploaded_) return false;
    for (int l = base_level_; l < depth_; ++l)
    {
        const double alpha = alpha_base * std::pow(4.0, l - base_level_);
        for (int nl = 0; nl < 2; ++nl)
        {
            const auto it = passes_.find(pass_key(l, fmm_mask_idx_gen(MASK_AMR2AMR, l - base_level_, nl), !laplace));
            if (it == passes_.end()) return false;
            for (const auto& m : it->second->mac)
                if (!has_ghat(laplace, m.d, alpha)) return false;
        }
    }
    return true;
}

static size_t round256(size_t b) { return (b + 255) / 256 * 256; }

size_t GpuSolver::region_bytes() const
{
    const size_t fb = round256(field_size() * sizeof(double));
    const size_t sb = round256(static_cast<size_t>(nb_) * nf_ * (precision_ == 0 ? sizeof(double2) : sizeof(float2)));
    return 4 * fb + 2 * sb;
}

void GpuSolver::solve_in_region(int src, int dst, bool laplace, double alpha_base, void* region, size_t bytes)
{
    if (!uploaded_) throw std::runtime_error("tree not uploaded");
    const size_t need = region_bytes();
    if (!region || bytes < need)
        throw std::length_error("solve needs " + std::to_string(need) + " bytes of work memory, the region has " +
                                std::to_string(bytes));
    if (!prepared(laplace, alpha_base))
        throw std::runtime_error("plans or kernel spectra of this solve are not built (prepare it first in this precision)");
    if (timer_.enabled()) throw std::runtime_error("kernel timing is on (it records events)");
    field_data(src);
    field_data(dst);
    IBLGF_HIP(hipSetDevice(device_));
    const size_t fb = round256(field_size() * sizeof(double));
    const size_t sb = round256(static_cast<size_t>(nb_) * nf_ * (precision_ == 0 ? sizeof(double2) : sizeof(float2)));
    char*        p = static_cast<char*>(region);
    // The region keeps whatever the last user left in it.
    noalloc_clear(stream_, region, need);
    const WorkPtrs saved = wp_;
    wp_ = WorkPtrs{reinterpret_cast<double*>(p), reinterpret_cast<double*>(p + fb),
        reinterpret_cast<double*>(p + 2 * fb), reinterpret_cast<double*>(p + 3 * fb), p + 4 * fb, p + 4 * fb + sb};
    try
    {
        if (laplace) lgf_batch_body(&src, &dst, 1);
        else
            if_batch_body(&src, &dst, 1, alpha_base);
    }
    catch (...)
    {
        wp_ = saved;
        throw;
    }
    wp_ = saved;
}

void GpuSolver::if_batch_body(const int* src, const int* dst, int k, double alpha_base)
{
    const Batch b = batch(k);
    double* st = wp_.src;
    double* tt = wp_.tgt;
    launch::zero_blocks(stream_, n_, st, p_nonleaf_.p, p_nonleaf_.size(), b);
    fill_zero(tt, b);
    launch::copy_blocks_table(
        stream_, n_, work_table(st, k), field_table(src, k), p_leaves_.p, p_leaves_.size(), false, k);
    for (int ls = depth_ - 2; ls >= base_level_; --ls)
        launch::transfer_gather(
            stream_, n_, st, st, p_coarsify_[ls].p, p_coarsify_[ls].size() / 9, m_cc_restrict_.p, true, b, false, mesh_);
    for (int l = depth_ - 2; l >= base_level_; --l)
    {
        launch::buffer_exchange(stream_, n_, st, p_level_slots_[l].p, p_level_slots_[l].size(), d_neighbor_.p, b);
        launch::transfer_scatter(
            stream_, n_, st, st, p_ccint_corr_[l].p, p_ccint_corr_[l].size() / 3, m_cc_interp_.p, false, b, mesh_);
    }
    for (int l = base_level_; l < depth_; ++l)
    {
        const double alpha = alpha_base * std::pow(4.0, l - base_level_);
        run_pass(pass(l, 0, true), 1.0, false, alpha, b);
        run_pass(pass(l, 1, true), 1.0, false, alpha, b);
        launch::copy_blocks_table(stream_, n_, field_table(dst, k), work_table(tt, k), p_level_slots_[l].p,
            p_level_slots_[l].size(), false, k);
    }
    IBLGF_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Immersed-boundary solves and level operators
// ---------------------------------------------------------------------------
void GpuSolver::apply_lgf_masked(int src, int dst, int mask_type)
{
    if (mask_type != MASK_IB2xIB && mask_type != MASK_xIB2IB && mask_type != MASK_STREAM)
        throw std::invalid_argument("mask type must be STREAM, IB2xIB or xIB2IB");
    check_batch(&src, &dst, 1);
    masked_body(src, dst, mask_type);
}

void GpuSolver::masked_body(int src, int dst, int mask_type)
{
    const Batch b = batch(1);
    // STREAM: one pass at the base level (poisson.hpp:609-611, 645-647); the
    // immersed-boundary types: one pass at the finest level.
    const int L = mask_type == MASK_STREAM ? base_level_ : depth_ - 1;
    double*   st = wp_.src;
    double*   tt = wp_.tgt;
    // The pass reads the source's level-L interiors only and adds into the
    // level-L targets only.
    launch::zero_blocks(stream_, n_, tt, p_level_slots_[L].p, p_level_slots_[L].size(), b);
    launch::copy_blocks(stream_, n_, st, F(src), p_level_slots_[L].p, p_level_slots_[L].size(), true);
    const double dx = dx_base_ / std::pow(2.0, L - base_level_);
    run_pass(pass_idx(L, fmm_mask_idx_gen(mask_type), false), dx * dx, true, 0.0, b);
    launch::copy_blocks(stream_, n_, F(dst), tt, p_leaf_level_[L].p, p_leaf_level_[L].size(), false);
    IBLGF_HIP(hipGetLastError());
}

// apply_lgf with fmm_type IB2AMR (poisson.hpp:613-616, 632-633, 645-650): the
// source's finest-level interiors coarsened to every level, then one FMM pass
// per level with the IB2AMR masks of that level (immersed-boundary blocks as
// sources, every block of the level a target), each level's leaves written.
void GpuSolver::apply_lgf_ib2amr(int src, int dst)
{
    check_batch(&src, &dst, 1);
    const Batch b = batch(1);
    const int   Lf = depth_ - 1;
    double*     st = wp_.src;
    double*     tt = wp_.tgt;
    fill_zero(st, b);
    fill_zero(tt, b);
    launch::copy_blocks(stream_, n_, st, F(src), p_level_slots_[Lf].p, p_level_slots_[Lf].size(), true);
    for (int ls = depth_ - 2; ls >= base_level_; --ls)
        launch::transfer_gather(
            stream_, n_, st, st, p_coarsify_[ls].p, p_coarsify_[ls].size() / 9, m_cc_restrict_.p, true, b, false, mesh_);
    for (int l = base_level_; l < depth_; ++l)
    {
        const double dx = dx_base_ / std::pow(2.0, l - base_level_);
        run_pass(pass_idx(l, fmm_mask_idx_gen(MASK_IB2AMR, l - base_level_), false), dx * dx, true, 0.0, b);
        launch::copy_blocks(stream_, n_, F(dst), tt, p_leaf_level_[l].p, p_leaf_level_[l].size(), false);
    }
    IBLGF_HIP(hipGetLastError());
}

void GpuSolver::apply_if_masked(int src, int dst, double alpha_base, int mask_type)
{
    apply_if_masked_batch(&src, &dst, 1, alpha_base, mask_type);
}

void GpuSolver::apply_if_masked_batch(const int* src, const int* dst, int k, double alpha_base, int mask_type)
{
    if (mask_type != MASK_IB2xIB && mask_type != MASK_xIB2IB)
        throw std::invalid_argument("mask type must be IB2xIB or xIB2IB");
    check_batch(src, dst, k);
    const Batch  b = batch(k);
    const int    L = depth_ - 1;
    double*      st = wp_.src;
    double*      tt = wp_.tgt;
    // The pass reads the source's level-L interiors only and adds into the
    // level-L targets only.
    launch::zero_blocks(stream_, n_, tt, p_level_slots_[L].p, p_level_slots_[L].size(), b);
    launch::copy_blocks_table(stream_, n_, work_table(st, k), field_table(src, k), p_level_slots_[L].p,
        p_level_slots_[L].size(), true, k);
    const double alpha = alpha_base * std::pow(4.0, L - base_level_);
    run_pass(pass_idx(L, fmm_mask_idx_gen(mask_type), true), 1.0, false, alpha, b);
    launch::copy_blocks_table(stream_, n_, field_table(dst, k), work_table(tt, k), p_level_slots_[L].p,
        p_level_slots_[L].size(), false, k);
    IBLGF_HIP(hipGetLastError());
}

void GpuSolver::block_op(int op, const int* src, const int* dst, int level, double scale)
{
    if (!uploaded_) throw std::runtime_error("tree not uploaded");
    if (level < 0 || level >= depth_) throw std::invalid_argument("level out of range");
    const int ncs = (op == OP_LAPLACE || op == OP_GRADIENT) ? 1 : 3;
    const int ncd = (op == OP_LAPLACE || op == OP_DIVERGENCE) ? 1 : 3;
    const double* s[3] = {nullptr, nullptr, nullptr};
    double*       d[3] = {nullptr, nullptr, nullptr};
    for (int c = 0; c < ncs; ++c) s[c] = F(src[c]);
    for (int c = 0; c < ncd; ++c) d[c] = F(dst[c]);
    for (int c = 0; c < ncs; ++c)
        for (int k = 0; k < ncd; ++k)
            if (src[c] == dst[k]) throw std::invalid_argument("block operator output aliases its input");
    const double dx = dx_base_ / std::pow(2.0, level - base_level_);
    double       fac = 1.0 / dx;
    if (op == OP_LAPLACE) fac = 1.0 / (dx * dx);
    if (op == OP_CURL_TRANSPOSE) fac = 1.0 / dx * scale;
    IBLGF_HIP(hipSetDevice(device_));
    launch::block_op(stream_, n_, op, s, d, p_level_slots_[level].p, p_level_slots_[level].size(), fac);
    IBLGF_HIP(hipGetLastError());
}

void GpuSolver::buffer_exchange(int f, int level)
{
    launch::buffer_exchange(stream_, n_, F(f), p_level_slots_[level].p, p_level_slots_[level].size(), d_neighbor_.p);
}

void GpuSolver::level_divergence(const int src[3], int dst, int level)
{
    for (int c = 0; c < 3; ++c) buffer_exchange(src[c], level);
    block_op(OP_DIVERGENCE, src, &dst, level, 1.0);
    clean_leaf_correction_boundary(dst, level, true, 2);
}

void GpuSolver::level_gradient(int src, const int dst[3], int level)
{
    block_op(OP_GRADIENT, &src, dst, level, 1.0);
    for (int c = 0; c < 3; ++c) buffer_exchange(dst[c], level);
}

// kind 0: leaf/correction boundary (neighbour missing, or a correction block
// when leaf_only); kind 1: IB region boundary (neighbour missing or not IB).
const DevArray<int>& GpuSolver::face_clean_list(int kind, int level, bool leaf_only)
{
    const int key = (level * 2 + kind) * 2 + (leaf_only ? 1 : 0);
    auto      it = clean_lists_.find(key);
    if (it != clean_lists_.end()) return it->second;
    static const int kFaceSlot[6] = {4, 10, 12, 14, 16, 22};
    std::vector<int> ent;
    for (int s = lvl_begin_[level]; s < lvl_begin_[level + 1]; ++s)
    {
        int mask = 0;
        for (int b = 0; b < 6; ++b)
        {
            const int nb = neighbor_[27 * s + kFaceSlot[b]];
            bool      clean = nb < 0;
            if (!clean && kind == 0) clean = leaf_only && (flags_[nb] & 2u);
            if (!clean && kind == 1) clean = !(flags_[nb] & 4u);
            if (clean) mask |= 1 << b;
        }
        if (mask) ent.insert(ent.end(), {s, mask});
    }
    return clean_lists_[key] = [&] {
        DevArray<int> a;
        a.upload(ent);
        return a;
    }();
}

void GpuSolver::clean_leaf_correction_boundary(int f, int level, bool leaf_only, int width)
{
    if (leaf_only)
    {
        const auto& c = corr_list(level);
        launch::zero_blocks(stream_, n_, F(f), c.p, c.size());
    }
    if (level == base_level_)
    {
        const auto& l = face_clean_list(0, level, leaf_only);
        launch::clean_faces(stream_, n_, F(f), l.p, l.size() / 2, width);
    }
}

const DevArray<int>& GpuSolver::corr_list(int level)
{
    auto it = corr_level_.find(level);
    if (it == corr_level_.end())
    {
        std::vector<int> cs;
        for (int s = lvl_begin_[level]; s < lvl_begin_[level + 1]; ++s)
            if (flags_[s] & 2u) cs.push_back(s);
        DevArray<int> a;
        a.upload(cs);
        it = corr_level_.emplace(level, std::move(a)).first;
    }
    return it->second;
}

void GpuSolver::clean_ib_region_boundary(int f, int level, int width)
{
    const auto& l = face_clean_list(1, level, false);
    launch::clean_faces(stream_, n_, F(f), l.p, l.size() / 2, width);
}

void GpuSolver::domain_clean(int f) { fill_zero(F(f)); }

void GpuSolver::add(int from, int to, double scale)
{
    launch::axpy_blocks(stream_, n_, F(to), F(from), p_all_.p, nb_, scale);
}

SolveCounts GpuSolver::lgf_counts() const
{
    SolveCounts c;
    c.n_leaf = n_leaf_;
    for (const auto& kv : passes_)
    {
        const PassPlan& p = *kv.second;
        if (p.neighbor_only || p.idx < 15 || p.idx >= 30) continue;  // Laplace AMR2AMR passes only
        c.passes += 1;
        c.n_src += p.fwd.size();
        c.n_tgt += p.inv.size();
        c.n_pair += p.npairs;
        c.n_transfer += p.ntransfer;
        for (const auto& m : p.mac)
        {
            c.n_entry += m.nentries;
            c.n_entry_full += m.nfull;
            c.n_pair_full += m.pairs_full;
            for (int i = 0; i < 8; ++i) c.masked_pop[i] += m.masked_pop[i];
        }
    }
    return c;
}

}  // namespace iblgf_amd
