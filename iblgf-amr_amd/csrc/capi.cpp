// C ABI of libiblgf_gpu.so: status codes around the host octree, the kernel
// tables and the device solver.  Declarations and reference citations are in
// include/iblgf_gpu.h.
#include "../../include/iblgf_gpu.h"

#include "gpu_solver.hpp"
#include "host/ib.hpp"
#include "host/octree.hpp"
#include "host/tables.hpp"
#include "ib_solver.hpp"
#include "learned_precond.hpp"
#include "ops.hpp"

#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>

using namespace iblgf_amd;

struct iblgf_tree
{
    Octree     tree;
    bool       masks_built = false;
    bool       subtract_non_leaf = false;  // the masks were built for it
    bool       has_ib = false;
    IbGeometry ib;
};

struct iblgf_gpu_ctx
{
    std::unique_ptr<GpuSolver> solver;
    std::unique_ptr<IbSystem>  ib;
    std::unique_ptr<LearnedPrecond> pc;
    int                        collect_stride = 0;
    CgStats                    last;
    DevArray<double>           ib_in, ib_out;
    std::string                error;
};

namespace {
thread_local std::string g_last_error;

int fail(int code, const std::string& msg, iblgf_gpu_ctx* ctx = nullptr)
{
    g_last_error = msg;
    if (ctx) ctx->error = msg;
    return code;
}

template<class F>
int guarded(iblgf_gpu_ctx* ctx, F&& f)
{
    try
    {
        f();
        return IBLGF_OK;
    }
    catch (const std::invalid_argument& e)
    {
        return fail(IBLGF_ERR_ARG, e.what(), ctx);
    }
    catch (const std::runtime_error& e)
    {
        const std::string m = e.what();
        if (m.rfind("HIP error", 0) == 0) return fail(IBLGF_ERR_DEVICE, m, ctx);
        if (m.find("unsupported") != std::string::npos) return fail(IBLGF_ERR_UNSUPPORTED, m, ctx);
        return fail(IBLGF_ERR_STATE, m, ctx);
    }
    catch (const std::exception& e)
    {
        return fail(IBLGF_ERR_INTERNAL, e.what(), ctx);
    }
    catch (...)
    {
        return fail(IBLGF_ERR_INTERNAL, "unknown error", ctx);
    }
}

const LaplaceLgfTable& shared_table()
{
    static std::once_flag  once;
    static LaplaceLgfTable tab;
    std::call_once(once, [] { tab.build(100); });
    return tab;
}
}  // namespace

extern "C" {

const char* iblgf_last_error(void) { return g_last_error.c_str(); }

int iblgf_tree_create(const int bd_base[3], const int bd_extent[3], int block_extent, double dx_base, int nregions,
    const int* regions, iblgf_tree** out)
{
    if (!bd_base || !bd_extent || !out || (nregions > 0 && !regions) || block_extent <= 0)
        return fail(IBLGF_ERR_ARG, "iblgf_tree_create: invalid argument");
    return guarded(nullptr, [&] {
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
        return fail(IBLGF_ERR_ARG, "iblgf_tree_import: invalid argument");
    return guarded(nullptr, [&] {
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
        return fail(IBLGF_ERR_ARG, "iblgf_tree_init_refine: invalid argument");
    return guarded(nullptr, [&] {
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
    if (!t) return fail(IBLGF_ERR_ARG, "iblgf_tree_build_masks: null tree");
    return guarded(nullptr, [&] {
        if (t->has_ib) update_ib_flags(t->tree, t->ib);
        t->tree.fmm_lgf_mask_build(subtract_non_leaf != 0);
        t->tree.fmm_ib_masks();
        t->masks_built = true;
        t->subtract_non_leaf = subtract_non_leaf != 0;
    });
}

int iblgf_tree_set_ib(iblgf_tree* t, const double* points, int npoints, int ib_level, int delta, double safety_dis)
{
    if (!t || !points || npoints <= 0 || ib_level < 0 || delta < IBLGF_DELTA_YANG3 || delta > IBLGF_DELTA_ROMA)
        return fail(IBLGF_ERR_ARG, "iblgf_tree_set_ib: invalid argument");
    return guarded(nullptr, [&] {
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
    if (!t || !t->has_ib || !sizes4) return fail(IBLGF_ERR_ARG, "iblgf_tree_ib_stencil: invalid argument");
    return guarded(nullptr, [&] {
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
    if (!t) return fail(IBLGF_ERR_ARG, "iblgf_tree_info: null tree");
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
    if (!t) return fail(IBLGF_ERR_ARG, "iblgf_tree_export: null tree");
    return guarded(nullptr, [&] {
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
    if (nmax < 0 || !g) return fail(IBLGF_ERR_ARG, "iblgf_lgf_table: invalid argument");
    return guarded(nullptr, [&] {
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
        return fail(IBLGF_ERR_ARG, "iblgf_lagrange_matrix: invalid argument");
    const auto A = relative_pos == 0 && pts_cap == 10 ? nli_matrix(n, 10) : cc_interp_matrix(n, pts_cap, relative_pos);
    std::copy(A.begin(), A.end(), out);
    return IBLGF_OK;
}

int iblgf_gpu_create(int device, iblgf_gpu_ctx** out)
{
    if (!out) return fail(IBLGF_ERR_ARG, "iblgf_gpu_create: null output");
    auto ctx = std::make_unique<iblgf_gpu_ctx>();
    const int rc = guarded(ctx.get(), [&] { ctx->solver = std::make_unique<GpuSolver>(device); });
    if (rc == IBLGF_OK) *out = ctx.release();
    return rc;
}

int iblgf_gpu_destroy(iblgf_gpu_ctx* ctx)
{
    if (!ctx) return fail(IBLGF_ERR_ARG, "iblgf_gpu_destroy: null context");
    const int rc = guarded(ctx, [&] {
        ctx->pc.reset();
        ctx->ib.reset();
        ctx->ib_in.release();
        ctx->ib_out.release();
        ctx->solver.reset();
    });
    delete ctx;
    return rc;
}

const char* iblgf_gpu_last_error(const iblgf_gpu_ctx* ctx) { return ctx ? ctx->error.c_str() : g_last_error.c_str(); }

#define NEED_CTX(name)                                                         \
    if (!ctx || !ctx->solver) return fail(IBLGF_ERR_ARG, name ": null context")

int iblgf_gpu_tree_upload(iblgf_gpu_ctx* ctx, const iblgf_tree* t)
{
    // Checked before the context: the refusal does not depend on the device.
    if (t && t->masks_built && t->subtract_non_leaf)
        return fail(IBLGF_ERR_UNSUPPORTED,
            "iblgf_gpu_tree_upload: masks built with subtract_non_leaf = 1; the device solves implement the "
            "reference default subtract_non_leaf = false (poisson.hpp:488-490, 670-716)",
            ctx);
    NEED_CTX("iblgf_gpu_tree_upload");
    if (!t) return fail(IBLGF_ERR_ARG, "iblgf_gpu_tree_upload: null tree", ctx);
    if (!t->masks_built) return fail(IBLGF_ERR_STATE, "iblgf_gpu_tree_upload: masks not built", ctx);
    return guarded(ctx, [&] {
        ctx->pc.reset();  // its fields and weights belong to the previous tree
        ctx->ib.reset();
        ctx->solver->upload_tree(t->tree);
    });
}

int iblgf_gpu_field_create(iblgf_gpu_ctx* ctx, int* field_id)
{
    NEED_CTX("iblgf_gpu_field_create");
    if (!field_id) return fail(IBLGF_ERR_ARG, "iblgf_gpu_field_create: null output", ctx);
    return guarded(ctx, [&] {
        if (ctx->solver->nblocks() == 0) throw std::runtime_error("tree not uploaded");
        *field_id = ctx->solver->field_create();
    });
}

int iblgf_gpu_field_device_ptr(iblgf_gpu_ctx* ctx, int field_id, void** dptr)
{
    NEED_CTX("iblgf_gpu_field_device_ptr");
    if (!dptr) return fail(IBLGF_ERR_ARG, "iblgf_gpu_field_device_ptr: null output", ctx);
    return guarded(ctx, [&] { *dptr = ctx->solver->field_ptr(field_id); });
}

int iblgf_gpu_field_upload(iblgf_gpu_ctx* ctx, int field_id, const double* host)
{
    NEED_CTX("iblgf_gpu_field_upload");
    if (!host) return fail(IBLGF_ERR_ARG, "iblgf_gpu_field_upload: null host pointer", ctx);
    return guarded(ctx, [&] { ctx->solver->field_upload(field_id, host); });
}

int iblgf_gpu_field_download(iblgf_gpu_ctx* ctx, int field_id, double* host)
{
    NEED_CTX("iblgf_gpu_field_download");
    if (!host) return fail(IBLGF_ERR_ARG, "iblgf_gpu_field_download: null host pointer", ctx);
    return guarded(ctx, [&] { ctx->solver->field_download(field_id, host); });
}

int iblgf_gpu_apply_lgf(iblgf_gpu_ctx* ctx, int src_field, int dst_field, int kernel, double alpha_base)
{
    NEED_CTX("iblgf_gpu_apply_lgf");
    if (kernel != IBLGF_KERNEL_LAPLACE && kernel != IBLGF_KERNEL_IF)
        return fail(IBLGF_ERR_ARG, "iblgf_gpu_apply_lgf: unknown kernel", ctx);
    if (src_field == dst_field) return fail(IBLGF_ERR_ARG, "iblgf_gpu_apply_lgf: src and dst must differ", ctx);
    return guarded(ctx, [&] {
        ctx->solver->field_ptr(src_field);
        ctx->solver->field_ptr(dst_field);
        if (kernel == IBLGF_KERNEL_LAPLACE) ctx->solver->apply_lgf(src_field, dst_field);
        else
            ctx->solver->apply_if(src_field, dst_field, alpha_base);
    });
}

int iblgf_gpu_apply_lgf_batch(iblgf_gpu_ctx* ctx, const int* src_fields, const int* dst_fields, int k, int kernel,
    double alpha_base)
{
    NEED_CTX("iblgf_gpu_apply_lgf_batch");
    if (kernel != IBLGF_KERNEL_LAPLACE && kernel != IBLGF_KERNEL_IF)
        return fail(IBLGF_ERR_ARG, "iblgf_gpu_apply_lgf_batch: unknown kernel", ctx);
    if (!src_fields || !dst_fields) return fail(IBLGF_ERR_ARG, "iblgf_gpu_apply_lgf_batch: null field list", ctx);
    if (k < 1 || k > iblgf_amd::GpuSolver::kMaxComponents)
        return fail(IBLGF_ERR_ARG, "iblgf_gpu_apply_lgf_batch: component count out of range", ctx);
    for (int c = 0; c < k; ++c)
        for (int e = 0; e < k; ++e)
            if (src_fields[c] == dst_fields[e] || (c != e && dst_fields[c] == dst_fields[e]))
                return fail(IBLGF_ERR_ARG, "iblgf_gpu_apply_lgf_batch: fields overlap", ctx);
    return guarded(ctx, [&] {
        if (kernel == IBLGF_KERNEL_LAPLACE) ctx->solver->apply_lgf_batch(src_fields, dst_fields, k);
        else
            ctx->solver->apply_if_batch(src_fields, dst_fields, k, alpha_base);
    });
}

int iblgf_gpu_set_graphs(iblgf_gpu_ctx* ctx, int on)
{
    NEED_CTX("iblgf_gpu_set_graphs");
    return guarded(ctx, [&] { ctx->solver->set_graphs(on != 0); });
}

int iblgf_gpu_set_concurrency(iblgf_gpu_ctx* ctx, int on)
{
    NEED_CTX("iblgf_gpu_set_concurrency");
    return guarded(ctx, [&] { ctx->solver->set_concurrency(on != 0); });
}

int iblgf_gpu_set_precision(iblgf_gpu_ctx* ctx, int precision)
{
    NEED_CTX("iblgf_gpu_set_precision");
    if (precision != IBLGF_PREC_FP64 && precision != IBLGF_PREC_MIXED)
        return fail(IBLGF_ERR_ARG, "iblgf_gpu_set_precision: unknown precision", ctx);
    return guarded(ctx, [&] { ctx->solver->set_precision(precision); });
}

int iblgf_gpu_prepare(iblgf_gpu_ctx* ctx, int kernel, double alpha_base)
{
    NEED_CTX("iblgf_gpu_prepare");
    return guarded(ctx, [&] {
        if (kernel == IBLGF_KERNEL_LAPLACE) ctx->solver->prepare_laplace();
        else if (kernel == IBLGF_KERNEL_IF)
            ctx->solver->prepare_if(alpha_base);
        else
            throw std::invalid_argument("unknown kernel");
    });
}

int iblgf_gpu_sync(iblgf_gpu_ctx* ctx)
{
    NEED_CTX("iblgf_gpu_sync");
    return guarded(ctx, [&] { ctx->solver->sync(); });
}

int iblgf_gpu_stream(iblgf_gpu_ctx* ctx, void** stream)
{
    NEED_CTX("iblgf_gpu_stream");
    if (!stream) return fail(IBLGF_ERR_ARG, "iblgf_gpu_stream: null output", ctx);
    *stream = reinterpret_cast<void*>(ctx->solver->stream());
    return IBLGF_OK;
}

int iblgf_gpu_timing_enable(iblgf_gpu_ctx* ctx, int enable)
{
    NEED_CTX("iblgf_gpu_timing_enable");
    return guarded(ctx, [&] {
        ctx->solver->sync();
        ctx->solver->timer().enable(enable != 0);
    });
}

int iblgf_gpu_timing_read(iblgf_gpu_ctx* ctx, double* ms5, long* launches5, long* work5)
{
    NEED_CTX("iblgf_gpu_timing_read");
    return guarded(ctx, [&] {
        ctx->solver->sync();
        auto& t = ctx->solver->timer();
        t.collect();
        for (int k = 0; k < KC_COUNT; ++k)
        {
            if (ms5) ms5[k] = t.total_ms[k];
            if (launches5) launches5[k] = t.launches[k];
            if (work5) work5[k] = t.work[k];
        }
    });
}

int iblgf_gpu_timing_reset(iblgf_gpu_ctx* ctx)
{
    NEED_CTX("iblgf_gpu_timing_reset");
    return guarded(ctx, [&] {
        ctx->solver->sync();
        ctx->solver->timer().collect();
        ctx->solver->timer().reset();
    });
}

int iblgf_gpu_block_op(iblgf_gpu_ctx* ctx, int op, const int* src_fields, const int* dst_fields, int level,
    double scale)
{
    NEED_CTX("iblgf_gpu_block_op");
    if (!src_fields || !dst_fields || op < IBLGF_OP_LAPLACE || op > IBLGF_OP_CURL_TRANSPOSE)
        return fail(IBLGF_ERR_ARG, "iblgf_gpu_block_op: invalid argument", ctx);
    return guarded(ctx, [&] { ctx->solver->block_op(op, src_fields, dst_fields, level, scale); });
}

int iblgf_gpu_buffer_exchange(iblgf_gpu_ctx* ctx, int field, int level)
{
    NEED_CTX("iblgf_gpu_buffer_exchange");
    return guarded(ctx, [&] {
        if (level < 0 || level >= ctx->solver->tree().depth) throw std::invalid_argument("level out of range");
        ctx->solver->field_ptr(field);
        ctx->solver->buffer_exchange(field, level);
    });
}

int iblgf_gpu_level_divergence(iblgf_gpu_ctx* ctx, const int src_fields[3], int dst_field, int level)
{
    NEED_CTX("iblgf_gpu_level_divergence");
    if (!src_fields) return fail(IBLGF_ERR_ARG, "iblgf_gpu_level_divergence: null fields", ctx);
    return guarded(ctx, [&] { ctx->solver->level_divergence(src_fields, dst_field, level); });
}

int iblgf_gpu_level_gradient(iblgf_gpu_ctx* ctx, int src_field, const int dst_fields[3], int level)
{
    NEED_CTX("iblgf_gpu_level_gradient");
    if (!dst_fields) return fail(IBLGF_ERR_ARG, "iblgf_gpu_level_gradient: null fields", ctx);
    return guarded(ctx, [&] { ctx->solver->level_gradient(src_field, dst_fields, level); });
}

int iblgf_gpu_clean_boundary(iblgf_gpu_ctx* ctx, int field, int level, int kind, int leaf_only, int width)
{
    NEED_CTX("iblgf_gpu_clean_boundary");
    if (kind != 0 && kind != 1) return fail(IBLGF_ERR_ARG, "iblgf_gpu_clean_boundary: unknown kind", ctx);
    return guarded(ctx, [&] {
        if (level < 0 || level >= ctx->solver->tree().depth) throw std::invalid_argument("level out of range");
        ctx->solver->field_ptr(field);
        if (kind == 0) ctx->solver->clean_leaf_correction_boundary(field, level, leaf_only != 0, width);
        else
            ctx->solver->clean_ib_region_boundary(field, level, width);
    });
}

int iblgf_gpu_apply_lgf_masked(iblgf_gpu_ctx* ctx, int src_field, int dst_field, int kernel, double alpha_base,
    int mask_type)
{
    NEED_CTX("iblgf_gpu_apply_lgf_masked");
    if (kernel != IBLGF_KERNEL_LAPLACE && kernel != IBLGF_KERNEL_IF)
        return fail(IBLGF_ERR_ARG, "iblgf_gpu_apply_lgf_masked: unknown kernel", ctx);
    if (mask_type != IBLGF_MASK_IB2xIB && mask_type != IBLGF_MASK_xIB2IB)
        return fail(IBLGF_ERR_ARG, "iblgf_gpu_apply_lgf_masked: unknown mask type", ctx);
    return guarded(ctx, [&] {
        ctx->solver->field_ptr(src_field);
        ctx->solver->field_ptr(dst_field);
        if (kernel == IBLGF_KERNEL_LAPLACE) ctx->solver->apply_lgf_masked(src_field, dst_field, mask_type);
        else
            ctx->solver->apply_if_masked(src_field, dst_field, alpha_base, mask_type);
    });
}

int iblgf_gpu_apply_lgf_typed(iblgf_gpu_ctx* ctx, int src_field, int dst_field, int kernel, double alpha_base,
    int fmm_type)
{
    // The type is checked first, so an unbuilt type is refused the same way
    // with or without a device.
    switch (fmm_type)
    {
        case IBLGF_FMM_AMR2AMR:
        case IBLGF_FMM_IB2xIB:
        case IBLGF_FMM_xIB2IB:
            break;
        case IBLGF_FMM_STREAM:
        case IBLGF_FMM_IB2AMR:
        case IBLGF_FMM_LAPLACIAN_BC:
        case IBLGF_FMM_REF_FOURIER_STREAM:
        case IBLGF_FMM_REF_FOURIER_LGF:
            return fail(IBLGF_ERR_UNSUPPORTED,
                "iblgf_gpu_apply_lgf_typed: fmm_type " + std::to_string(fmm_type) +
                    " has no device path (built: AMR2AMR, IB2xIB, xIB2IB)",
                ctx);
        default:
            return fail(IBLGF_ERR_ARG, "iblgf_gpu_apply_lgf_typed: unknown fmm_type", ctx);
    }
    if (fmm_type == IBLGF_FMM_AMR2AMR) return iblgf_gpu_apply_lgf(ctx, src_field, dst_field, kernel, alpha_base);
    return iblgf_gpu_apply_lgf_masked(ctx, src_field, dst_field, kernel, alpha_base, fmm_type);
}

int iblgf_gpu_ib_create(iblgf_gpu_ctx* ctx, const iblgf_tree* t)
{
    NEED_CTX("iblgf_gpu_ib_create");
    if (!t || !t->has_ib) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ib_create: tree without immersed-boundary points", ctx);
    return guarded(ctx, [&] {
        ctx->pc.reset();
        ctx->ib.reset();
        ctx->ib = std::make_unique<IbSystem>(*ctx->solver, t->ib);
        ctx->ib_in.alloc(3 * static_cast<size_t>(t->ib.size()));
        ctx->ib_out.alloc(3 * static_cast<size_t>(t->ib.size()));
    });
}

#define NEED_IB(name)                                                                              \
    NEED_CTX(name);                                                                                \
    if (!ctx->ib) return fail(IBLGF_ERR_STATE, name ": no immersed-boundary system", ctx)

int iblgf_gpu_ib_info(iblgf_gpu_ctx* ctx, int* npoints, int* nowned, long* nnz)
{
    NEED_IB("iblgf_gpu_ib_info");
    const auto& st = ctx->ib->stencil();
    if (npoints) *npoints = st.npoints;
    if (nowned)
    {
        int c = 0;
        for (int o : st.owned) c += o;
        *nowned = c;
    }
    if (nnz) *nnz = static_cast<long>(st.interp_w.size());
    return IBLGF_OK;
}

namespace {
void ib_to_device(iblgf_gpu_ctx* ctx, DevArray<double>& d, const double* h)
{
    const hipError_t e = hipMemcpyAsync(d.p, h, d.n * sizeof(double), hipMemcpyHostToDevice, ctx->solver->stream());
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e));
}
void ib_to_host(iblgf_gpu_ctx* ctx, double* h, const DevArray<double>& d)
{
    hipError_t e = hipMemcpyAsync(h, d.p, d.n * sizeof(double), hipMemcpyDeviceToHost, ctx->solver->stream());
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->solver->stream());
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e));
}
}  // namespace

int iblgf_gpu_ib_smear(iblgf_gpu_ctx* ctx, const double* force, const int fields[3])
{
    NEED_IB("iblgf_gpu_ib_smear");
    if (!force || !fields) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ib_smear: null argument", ctx);
    return guarded(ctx, [&] {
        for (int c = 0; c < 3; ++c) ctx->solver->field_ptr(fields[c]);
        ib_to_device(ctx, ctx->ib_in, force);
        ctx->ib->smear(ctx->ib_in.p, fields);
        ctx->solver->sync();
    });
}

int iblgf_gpu_ib_project(iblgf_gpu_ctx* ctx, const int fields[3], double* force)
{
    NEED_IB("iblgf_gpu_ib_project");
    if (!force || !fields) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ib_project: null argument", ctx);
    return guarded(ctx, [&] {
        for (int c = 0; c < 3; ++c) ctx->solver->field_ptr(fields[c]);
        ctx->ib->project(fields, ctx->ib_out.p);
        ib_to_host(ctx, force, ctx->ib_out);
    });
}

int iblgf_gpu_ib_apply(iblgf_gpu_ctx* ctx, const double* fin, double* fout, double alpha)
{
    NEED_IB("iblgf_gpu_ib_apply");
    if (!fin || !fout) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ib_apply: null argument", ctx);
    return guarded(ctx, [&] {
        ib_to_device(ctx, ctx->ib_in, fin);
        ctx->ib->apply(ctx->ib_in.p, ctx->ib_out.p, alpha);
        ib_to_host(ctx, fout, ctx->ib_out);
    });
}

namespace {
int ib_solve(iblgf_gpu_ctx* ctx, const double* uc, double* f, double alpha, double tol, int maxit, int use_mlp,
    bool flexible, int* iterations, double* residual, double* history)
{
    return guarded(ctx, [&] {
        if (use_mlp && !ctx->pc) throw std::runtime_error("no learned preconditioner");
        DevArray<double> d_uc;
        d_uc.alloc(ctx->ib_in.n);
        ib_to_device(ctx, d_uc, uc);
        ib_to_device(ctx, ctx->ib_out, f);
        const bool collect = ctx->pc && ctx->collect_stride > 0;
        if (collect) ctx->pc->begin_collect(ctx->collect_stride);
        try
        {
            if (!flexible) ctx->last = ctx->ib->cg(d_uc.p, ctx->ib_out.p, alpha, tol, maxit);
            else if (use_mlp)
            {
                LearnedPrecond* pc = ctx->pc.get();
                FcgMonitor      mon;
                mon.enabled = true;
                ctx->last = ctx->ib->fcg(d_uc.p, ctx->ib_out.p, alpha, tol, maxit,
                    [pc](hipStream_t s, const double* r, double* z) { pc->apply(s, r, z); }, mon);
            }
            else
                ctx->last = ctx->ib->fcg(d_uc.p, ctx->ib_out.p, alpha, tol, maxit, Preconditioner());
        }
        catch (...)
        {
            if (collect) ctx->ib->on_iteration = nullptr;
            throw;
        }
        if (collect) ctx->pc->end_collect(ctx->ib_out.p);
        ib_to_host(ctx, f, ctx->ib_out);
        if (iterations) *iterations = ctx->last.iterations;
        if (residual) *residual = ctx->last.residual;
        if (history)
            for (size_t k = 0; k < ctx->last.history.size(); ++k) history[k] = ctx->last.history[k];
    });
}
}  // namespace

int iblgf_gpu_ib_cg(iblgf_gpu_ctx* ctx, const double* uc, double* f, double alpha, double tol, int maxit,
    int* iterations, double* residual, double* history)
{
    NEED_IB("iblgf_gpu_ib_cg");
    if (!uc || !f || maxit < 0 || !(tol > 0.0)) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ib_cg: invalid argument", ctx);
    return ib_solve(ctx, uc, f, alpha, tol, maxit, 0, false, iterations, residual, history);
}

int iblgf_gpu_ib_fcg(iblgf_gpu_ctx* ctx, const double* uc, double* f, double alpha, double tol, int maxit,
    int use_mlp, int* iterations, double* residual, double* history)
{
    NEED_IB("iblgf_gpu_ib_fcg");
    if (!uc || !f || maxit < 0 || !(tol > 0.0)) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ib_fcg: invalid argument", ctx);
    return ib_solve(ctx, uc, f, alpha, tol, maxit, use_mlp, true, iterations, residual, history);
}

int iblgf_gpu_ib_stats(iblgf_gpu_ctx* ctx, double* seconds, int* applies, int* fell_back, int* stop)
{
    NEED_IB("iblgf_gpu_ib_stats");
    if (seconds) *seconds = ctx->last.seconds;
    if (applies) *applies = ctx->last.applies;
    if (fell_back) *fell_back = ctx->last.fell_back ? 1 : 0;
    if (stop) *stop = ctx->last.stop == "converged" ? 0 : (ctx->last.stop == "iteration limit" ? 1 : 2);
    return IBLGF_OK;
}

int iblgf_gpu_ib_switch_iteration(iblgf_gpu_ctx* ctx, int* iteration)
{
    NEED_IB("iblgf_gpu_ib_switch_iteration");
    if (!iteration) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ib_switch_iteration: null output", ctx);
    *iteration = ctx->last.switched_at;
    return IBLGF_OK;
}

int iblgf_gpu_ib_recycle(iblgf_gpu_ctx* ctx, int capacity)
{
    NEED_IB("iblgf_gpu_ib_recycle");
    if (capacity < 0 || capacity > 4096) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ib_recycle: capacity out of range", ctx);
    return guarded(ctx, [&] { ctx->ib->set_recycling(capacity); });
}

int iblgf_gpu_ib_precond_create(iblgf_gpu_ctx* ctx, const iblgf_tree* t, int K, int hidden, uint64_t seed)
{
    NEED_IB("iblgf_gpu_ib_precond_create");
    if (!t || !t->has_ib) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ib_precond_create: tree without points", ctx);
    return guarded(ctx, [&] {
        ctx->pc.reset();
        ctx->collect_stride = 0;
        ctx->pc = std::make_unique<LearnedPrecond>(*ctx->ib, t->ib, K, hidden, seed);
    });
}

#define NEED_PC(name)                                                                              \
    NEED_IB(name);                                                                                 \
    if (!ctx->pc) return fail(IBLGF_ERR_STATE, name ": no learned preconditioner", ctx)

int iblgf_gpu_ib_precond_collect(iblgf_gpu_ctx* ctx, int stride)
{
    NEED_PC("iblgf_gpu_ib_precond_collect");
    if (stride < 0) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ib_precond_collect: negative stride", ctx);
    ctx->collect_stride = stride;
    return IBLGF_OK;
}

int iblgf_gpu_ib_precond_train(iblgf_gpu_ctx* ctx, int steps, int batch_rows, double lr, double* loss, int* npairs)
{
    NEED_PC("iblgf_gpu_ib_precond_train");
    if (steps < 0 || batch_rows < 1 || !(lr > 0)) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ib_precond_train: invalid argument", ctx);
    return guarded(ctx, [&] {
        const double l = steps > 0 ? ctx->pc->train(steps, batch_rows, (float)lr) : 0.0;
        if (loss) *loss = l;
        if (npairs) *npairs = ctx->pc->npairs();
    });
}

int iblgf_gpu_ib_precond_info(iblgf_gpu_ctx* ctx, int* K, int* hidden, int* dim, long* nparam, double* out_scale)
{
    NEED_PC("iblgf_gpu_ib_precond_info");
    if (K) *K = ctx->pc->feature_dim() / 6;
    if (hidden) *hidden = ctx->pc->hidden();
    if (dim) *dim = ctx->pc->feature_dim();
    if (nparam) *nparam = static_cast<long>(ctx->pc->params().size());
    if (out_scale) *out_scale = ctx->pc->output_scale();
    return IBLGF_OK;
}

int iblgf_gpu_ib_precond_params(iblgf_gpu_ctx* ctx, float* params, long nparam)
{
    NEED_PC("iblgf_gpu_ib_precond_params");
    if (!params) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ib_precond_params: null output", ctx);
    return guarded(ctx, [&] {
        const auto p = ctx->pc->params();
        if ((long)p.size() != nparam) throw std::invalid_argument("parameter count mismatch");
        std::copy(p.begin(), p.end(), params);
    });
}

int iblgf_gpu_ib_precond_set_params(iblgf_gpu_ctx* ctx, const float* params, long nparam)
{
    NEED_PC("iblgf_gpu_ib_precond_set_params");
    if (!params || nparam <= 0) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ib_precond_set_params: invalid argument", ctx);
    return guarded(ctx, [&] { ctx->pc->set_params(std::vector<float>(params, params + nparam)); });
}

int iblgf_gpu_ib_precond_features(iblgf_gpu_ctx* ctx, const double* r, float* X)
{
    NEED_PC("iblgf_gpu_ib_precond_features");
    if (!r || !X) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ib_precond_features: null argument", ctx);
    return guarded(ctx, [&] {
        ib_to_device(ctx, ctx->ib_in, r);
        const auto v = ctx->pc->features(ctx->ib_in.p);
        std::copy(v.begin(), v.end(), X);
    });
}

int iblgf_gpu_ib_precond_grad(iblgf_gpu_ctx* ctx, const float* X, const float* T, int rows, float* grads, double* loss)
{
    NEED_PC("iblgf_gpu_ib_precond_grad");
    if (!X || !T || !grads || rows < 1) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ib_precond_grad: invalid argument", ctx);
    return guarded(ctx, [&] {
        const int          D = ctx->pc->feature_dim();
        std::vector<float> x(X, X + (size_t)rows * D), t(T, T + 3 * (size_t)rows), g;
        const double       l = ctx->pc->loss_and_grad(x, t, rows, g);
        std::copy(g.begin(), g.end(), grads);
        if (loss) *loss = l;
    });
}

int iblgf_gpu_counts(iblgf_gpu_ctx* ctx, long* c)
{
    NEED_CTX("iblgf_gpu_counts");
    if (!c) return fail(IBLGF_ERR_ARG, "iblgf_gpu_counts: null output", ctx);
    const SolveCounts s = ctx->solver->lgf_counts();
    c[0] = s.passes;
    c[1] = s.n_src;
    c[2] = s.n_tgt;
    c[3] = s.n_pair;
    c[4] = s.n_transfer;
    c[5] = s.n_leaf;
    return IBLGF_OK;
}

int iblgf_gpu_mac_counts(iblgf_gpu_ctx* ctx, long* c)
{
    NEED_CTX("iblgf_gpu_mac_counts");
    if (!c) return fail(IBLGF_ERR_ARG, "iblgf_gpu_mac_counts: null output", ctx);
    const SolveCounts s = ctx->solver->lgf_counts();
    c[0] = s.n_entry;
    c[1] = s.n_entry_full;
    c[2] = s.n_pair_full;
    c[3] = s.n_pair;
    for (int i = 0; i < 8; ++i) c[4 + i] = s.masked_pop[i];
    return IBLGF_OK;
}

}  // extern "C"
