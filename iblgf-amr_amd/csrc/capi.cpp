// Device half of the C ABI of libiblgf_gpu.so: contexts, fields, solves,
// operators and the immersed-boundary system.  Declarations and reference
// citations are in include/iblgf_gpu.h.
#include "capi_common.hpp"

#include "gpu_solver.hpp"
#include "host/chombo.hpp"
#include "ib_solver.hpp"
#include "learned_precond.hpp"
#include "noalloc.hpp"
#include "ops.hpp"

#include <algorithm>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

using namespace iblgf_amd;

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
int fail(int code, const std::string& msg, iblgf_gpu_ctx* ctx = nullptr)
{
    return capi::fail(code, msg, ctx ? &ctx->error : nullptr);
}

template<class F>
int guarded(iblgf_gpu_ctx* ctx, F&& f)
{
    return capi::guarded(ctx ? &ctx->error : nullptr, std::forward<F>(f));
}
}  // namespace

extern "C" {

int iblgf_gpu_create(int device, iblgf_gpu_ctx** out)
{
    if (!out) return fail(IBLGF_ERR_ARG, "iblgf_gpu_create: null output");
    auto ctx = std::make_unique<iblgf_gpu_ctx>();
    const int rc = guarded(ctx.get(), [&] {
        ctx->solver = std::make_unique<GpuSolver>(device);  // makes `device` current
        noalloc_scratch_init(device);
    });
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

const char* iblgf_gpu_last_error(const iblgf_gpu_ctx* ctx) { return ctx ? ctx->error.c_str() : capi::last_error().c_str(); }

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

int iblgf_gpu_tree_readapt(iblgf_gpu_ctx* ctx, const iblgf_tree* t, int nf, const int* fields, const int* mesh,
    const int* comp, int* nparents)
{
    NEED_CTX("iblgf_gpu_tree_readapt");
    if (!t) return fail(IBLGF_ERR_ARG, "iblgf_gpu_tree_readapt: null tree", ctx);
    if (!t->masks_built) return fail(IBLGF_ERR_STATE, "iblgf_gpu_tree_readapt: masks not built", ctx);
    if (t->subtract_non_leaf)
        return fail(IBLGF_ERR_UNSUPPORTED, "iblgf_gpu_tree_readapt: masks built with subtract_non_leaf = 1", ctx);
    if (nf < 0 || (nf > 0 && (!fields || !mesh || !comp)))
        return fail(IBLGF_ERR_ARG, "iblgf_gpu_tree_readapt: bad field list", ctx);
    return guarded(ctx, [&] {
        ctx->pc.reset();
        ctx->ib.reset();
        const int np = ctx->solver->readapt(t->tree, fields, mesh, comp, nf);
        if (nparents) *nparents = np;
    });
}

int iblgf_gpu_coarsen(iblgf_gpu_ctx* ctx, const int* fields, int k, int mesh_type)
{
    NEED_CTX("iblgf_gpu_coarsen");
    if (!fields || k < 1 || k > GpuSolver::kMaxComponents)
        return fail(IBLGF_ERR_ARG, "iblgf_gpu_coarsen: bad field list", ctx);
    return guarded(ctx, [&] {
        for (int c = 0; c < k; ++c) ctx->solver->field_ptr(fields[c]);
        ctx->solver->coarsen(fields, k, mesh_type);
    });
}

int iblgf_gpu_forest_upload(iblgf_gpu_ctx* ctx, int ntrees, const iblgf_tree* const* trees)
{
    NEED_CTX("iblgf_gpu_forest_upload");
    if (ntrees < 1 || !trees) return fail(IBLGF_ERR_ARG, "iblgf_gpu_forest_upload: no trees", ctx);
    std::vector<const Octree*> ts;
    for (int i = 0; i < ntrees; ++i)
    {
        const iblgf_tree* t = trees[i];
        if (!t) return fail(IBLGF_ERR_ARG, "iblgf_gpu_forest_upload: null tree", ctx);
        if (!t->masks_built) return fail(IBLGF_ERR_STATE, "iblgf_gpu_forest_upload: masks not built", ctx);
        if (t->subtract_non_leaf)
            return fail(IBLGF_ERR_UNSUPPORTED, "iblgf_gpu_forest_upload: masks built with subtract_non_leaf = 1", ctx);
        ts.push_back(&t->tree);
    }
    return guarded(ctx, [&] {
        ctx->pc.reset();
        ctx->ib.reset();
        ctx->solver->upload_forest(ts);
    });
}

int iblgf_gpu_forest_slots(iblgf_gpu_ctx* ctx, int tree, int* slots)
{
    NEED_CTX("iblgf_gpu_forest_slots");
    if (!slots) return fail(IBLGF_ERR_ARG, "iblgf_gpu_forest_slots: null output", ctx);
    return guarded(ctx, [&] {
        GpuSolver& s = *ctx->solver;
        if (tree < 0 || tree >= s.num_trees()) throw std::invalid_argument("iblgf_gpu_forest_slots: tree index");
        // the tree's own slot order: level-major, key-sorted (as iblgf_tree_export)
        const std::vector<int>& so = s.forest_slot_of(tree);
        std::vector<std::pair<int, int>> own;  // (combined slot, ...) of the tree's octants with data
        for (int o = 0; o < (int)so.size(); ++o)
            if (so[o] >= 0) own.emplace_back(so[o], o);
        // combined slots keep each tree's level-major key order, so sorting them
        // gives the tree's own order
        std::sort(own.begin(), own.end());
        for (size_t k = 0; k < own.size(); ++k) slots[k] = own[k].first;
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
    return guarded(ctx, [&] { *dptr = ctx->solver->field_data(field_id); });
}

int iblgf_gpu_field_create_f32(iblgf_gpu_ctx* ctx, int* field_id)
{
    NEED_CTX("iblgf_gpu_field_create_f32");
    if (!field_id) return fail(IBLGF_ERR_ARG, "iblgf_gpu_field_create_f32: null output", ctx);
    return guarded(ctx, [&] {
        if (ctx->solver->nblocks() == 0) throw std::runtime_error("tree not uploaded");
        *field_id = ctx->solver->field_create_f32();
    });
}

int iblgf_gpu_field_upload_f32(iblgf_gpu_ctx* ctx, int field_id, const float* host)
{
    NEED_CTX("iblgf_gpu_field_upload_f32");
    if (!host) return fail(IBLGF_ERR_ARG, "iblgf_gpu_field_upload_f32: null host pointer", ctx);
    return guarded(ctx, [&] { ctx->solver->field_upload_f32(field_id, host); });
}

int iblgf_gpu_field_download_f32(iblgf_gpu_ctx* ctx, int field_id, float* host)
{
    NEED_CTX("iblgf_gpu_field_download_f32");
    if (!host) return fail(IBLGF_ERR_ARG, "iblgf_gpu_field_download_f32: null host pointer", ctx);
    return guarded(ctx, [&] { ctx->solver->field_download_f32(field_id, host); });
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

int iblgf_gpu_chombo_write(iblgf_gpu_ctx* ctx, const char* path, int nfields, const char* const* names,
    const int* ncomp, const int* field_ids, int include_correction, int base_correction_only, double time, double dt)
{
    NEED_CTX("iblgf_gpu_chombo_write");
    if (!path || nfields < 1 || !names || !ncomp || !field_ids)
        return fail(IBLGF_ERR_ARG, "iblgf_gpu_chombo_write: invalid argument", ctx);
    return guarded(ctx, [&] {
        GpuSolver& s = *ctx->solver;
        if (s.num_trees() != 1) throw capi::Unsupported("iblgf_gpu_chombo_write: a forest context holds several trees");
        const size_t per = s.field_size();
        std::vector<std::vector<double>> host(nfields);
        std::vector<ChomboField>         f;
        int                              id = 0;
        for (int i = 0; i < nfields; ++i)
        {
            if (!names[i] || ncomp[i] < 1) throw std::invalid_argument("iblgf_gpu_chombo_write: bad field");
            host[i].resize(per * ncomp[i]);
            for (int c = 0; c < ncomp[i]; ++c) s.field_download(field_ids[id++], host[i].data() + per * c);
            f.push_back(ChomboField{names[i], ncomp[i], host[i].data()});
        }
        chombo_write(s.tree(), path, f, include_correction != 0, base_correction_only != 0, time, dt);
    });
}

int iblgf_gpu_chombo_read(iblgf_gpu_ctx* ctx, const char* path, const char* name, int ncomp, int mesh_type,
    const int* field_ids)
{
    NEED_CTX("iblgf_gpu_chombo_read");
    if (!path || !name || ncomp < 1 || mesh_type < 0 || mesh_type > 3 || !field_ids)
        return fail(IBLGF_ERR_ARG, "iblgf_gpu_chombo_read: invalid argument", ctx);
    return guarded(ctx, [&] {
        GpuSolver& s = *ctx->solver;
        if (s.num_trees() != 1) throw capi::Unsupported("iblgf_gpu_chombo_read: a forest context holds several trees");
        const size_t        per = s.field_size();
        std::vector<double> host(per * ncomp);
        for (int c = 0; c < ncomp; ++c) s.field_download(field_ids[c], host.data() + per * c);
        chombo_read(s.tree(), path, name, ncomp, mesh_type, host.data());
        for (int c = 0; c < ncomp; ++c) s.field_upload(field_ids[c], host.data() + per * c);
    });
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
    return iblgf_gpu_apply_lgf_mesh(ctx, src_fields, dst_fields, k, kernel, alpha_base, IBLGF_MESH_CELL);
}

int iblgf_gpu_apply_lgf_mesh(iblgf_gpu_ctx* ctx, const int* src_fields, const int* dst_fields, int k, int kernel,
    double alpha_base, int mesh_type)
{
    NEED_CTX("iblgf_gpu_apply_lgf_batch");
    if (mesh_type != IBLGF_MESH_CELL && mesh_type != IBLGF_MESH_FACE && mesh_type != IBLGF_MESH_EDGE)
        return fail(IBLGF_ERR_ARG, "iblgf_gpu_apply_lgf_mesh: unknown mesh type", ctx);
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
        if (kernel == IBLGF_KERNEL_LAPLACE) ctx->solver->apply_lgf_batch(src_fields, dst_fields, k, mesh_type);
        else
            ctx->solver->apply_if_batch(src_fields, dst_fields, k, alpha_base, mesh_type);
    });
}

int iblgf_gpu_solve_sequence(iblgf_gpu_ctx* ctx, int nitems, const iblgf_solve_item* items)
{
    NEED_CTX("iblgf_gpu_solve_sequence");
    if (nitems < 1 || !items) return fail(IBLGF_ERR_ARG, "iblgf_gpu_solve_sequence: no items", ctx);
    std::vector<GpuSolver::SolveItem> v(nitems);
    for (int i = 0; i < nitems; ++i)
    {
        const iblgf_solve_item& it = items[i];
        if ((it.kernel != IBLGF_KERNEL_LAPLACE && it.kernel != IBLGF_KERNEL_IF) || it.k < 1 ||
            it.k > GpuSolver::kMaxComponents || !it.src_fields || !it.dst_fields)
            return fail(IBLGF_ERR_ARG, "iblgf_gpu_solve_sequence: item " + std::to_string(i) + " is invalid", ctx);
        if (it.mesh_type != IBLGF_MESH_CELL && it.mesh_type != IBLGF_MESH_FACE && it.mesh_type != IBLGF_MESH_EDGE)
            return fail(IBLGF_ERR_ARG, "iblgf_gpu_solve_sequence: item " + std::to_string(i) + " mesh type", ctx);
        v[i].laplace = it.kernel == IBLGF_KERNEL_LAPLACE;
        v[i].alpha = it.alpha_base;
        v[i].mesh = it.mesh_type;
        v[i].src.assign(it.src_fields, it.src_fields + it.k);
        v[i].dst.assign(it.dst_fields, it.dst_fields + it.k);
    }
    return guarded(ctx, [&] { ctx->solver->solve_sequence(v); });
}

int iblgf_gpu_autotune(iblgf_gpu_ctx* ctx, const int* src_fields, const int* dst_fields, int k, int kernel,
    double alpha_base, int reps, char* report, int report_len)
{
    NEED_CTX("iblgf_gpu_autotune");
    if (!src_fields || !dst_fields || k < 1 || k > GpuSolver::kMaxComponents ||
        (kernel != IBLGF_KERNEL_LAPLACE && kernel != IBLGF_KERNEL_IF) || reps < 1 || (report_len > 0 && !report))
        return fail(IBLGF_ERR_ARG, "iblgf_gpu_autotune: invalid argument", ctx);
    return guarded(ctx, [&] {
        const std::string r =
            ctx->solver->autotune(src_fields, dst_fields, k, kernel == IBLGF_KERNEL_LAPLACE, alpha_base, reps);
        if (report_len > 0)
        {
            const size_t n = std::min(r.size(), static_cast<size_t>(report_len - 1));
            std::copy(r.begin(), r.begin() + static_cast<long>(n), report);
            report[n] = '\0';
        }
    });
}

int iblgf_gpu_set_launch_shapes(iblgf_gpu_ctx* ctx, int mac_policy, int dft_grid)
{
    NEED_CTX("iblgf_gpu_set_launch_shapes");
    if (mac_policy < 0 || mac_policy > 2 || dft_grid < 0 || dft_grid > 1)
        return fail(IBLGF_ERR_ARG, "iblgf_gpu_set_launch_shapes: shape out of range", ctx);
    return guarded(ctx, [&] { ctx->solver->set_launch_shapes(mac_policy, dft_grid); });
}

int iblgf_gpu_launch_shapes(iblgf_gpu_ctx* ctx, int* mac_policy, int* dft_grid, int* streams)
{
    NEED_CTX("iblgf_gpu_launch_shapes");
    if (!mac_policy || !dft_grid || !streams) return fail(IBLGF_ERR_ARG, "iblgf_gpu_launch_shapes: null output", ctx);
    *mac_policy = ctx->solver->mac_policy();
    *dft_grid = ctx->solver->dft_grid();
    *streams = ctx->solver->concurrency() ? 2 : 1;
    return IBLGF_OK;
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
    if (precision < IBLGF_PREC_FP64 || precision > IBLGF_PREC_FP16X3_FORWARD)
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

int iblgf_gpu_smooth_boundary(iblgf_gpu_ctx* ctx, int field, int level, int leaf_only)
{
    NEED_CTX("iblgf_gpu_smooth_boundary");
    return guarded(ctx, [&] {
        if (level < 0 || level >= ctx->solver->tree().depth) throw std::invalid_argument("level out of range");
        ctx->solver->field_ptr(field);
        ctx->solver->smooth_boundary(field, level, leaf_only != 0);
    });
}

int iblgf_gpu_block_norms(iblgf_gpu_ctx* ctx, const int* fields, int k, double* maxnorm, double* rms)
{
    NEED_CTX("iblgf_gpu_block_norms");
    if (!fields || k < 1 || k > 3) return fail(IBLGF_ERR_ARG, "iblgf_gpu_block_norms: 1 to 3 fields", ctx);
    return guarded(ctx, [&] {
        if (ctx->solver->nblocks() == 0) throw std::runtime_error("tree not uploaded");
        ctx->solver->block_norms(fields, k, maxnorm, rms);
    });
}

int iblgf_gpu_up_and_down(iblgf_gpu_ctx* ctx, const int* fields, int k, int mesh_type)
{
    NEED_CTX("iblgf_gpu_up_and_down");
    if (!fields || k < 1 || k > GpuSolver::kMaxComponents)
        return fail(IBLGF_ERR_ARG, "iblgf_gpu_up_and_down: bad field list", ctx);
    if (mesh_type != IBLGF_MESH_CELL && mesh_type != IBLGF_MESH_FACE && mesh_type != IBLGF_MESH_EDGE)
        return fail(IBLGF_ERR_ARG, "iblgf_gpu_up_and_down: unknown mesh type", ctx);
    return guarded(ctx, [&] {
        for (int c = 0; c < k; ++c) ctx->solver->field_ptr(fields[c]);
        ctx->solver->up_and_down(fields, k, mesh_type);
    });
}

int iblgf_gpu_nonlinear(iblgf_gpu_ctx* ctx, const int u_fields[3], const int w_fields[3], const int dst_fields[3],
    int level, double scale)
{
    NEED_CTX("iblgf_gpu_nonlinear");
    if (!u_fields || !w_fields || !dst_fields) return fail(IBLGF_ERR_ARG, "iblgf_gpu_nonlinear: null fields", ctx);
    for (int c = 0; c < 3; ++c)
        for (int e = 0; e < 3; ++e)
            if (dst_fields[c] == u_fields[e] || dst_fields[c] == w_fields[e] || (c != e && dst_fields[c] == dst_fields[e]))
                return fail(IBLGF_ERR_ARG, "iblgf_gpu_nonlinear: fields overlap", ctx);
    return guarded(ctx, [&] {
        if (level < 0 || level >= ctx->solver->tree().depth) throw std::invalid_argument("level out of range");
        for (int c = 0; c < 3; ++c)
        {
            ctx->solver->field_ptr(u_fields[c]);
            ctx->solver->field_ptr(w_fields[c]);
            ctx->solver->field_ptr(dst_fields[c]);
        }
        ctx->solver->nonlinear(u_fields, w_fields, dst_fields, level, scale);
    });
}

int iblgf_gpu_ifherk_step(iblgf_gpu_ctx* ctx, const int u_fields[3], int p_field, double dt, double Re,
    const double* frame_vel, int refresh)
{
    NEED_CTX("iblgf_gpu_ifherk_step");
    if (!u_fields) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ifherk_step: null velocity fields", ctx);
    if (refresh < 0 || refresh > 2) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ifherk_step: refresh must be 0, 1 or 2", ctx);
    if (!(dt > 0.0) || !(Re > 0.0)) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ifherk_step: dt and Re must be positive", ctx);
    return guarded(ctx, [&] {
        for (int c = 0; c < 3; ++c) ctx->solver->field_data(u_fields[c]);
        if (p_field >= 0) ctx->solver->field_data(p_field);
        ctx->solver->ifherk_step(u_fields, p_field, dt, Re, frame_vel, refresh);
    });
}

int iblgf_gpu_ifherk_step_ib(iblgf_gpu_ctx* ctx, const int u_fields[3], int p_field, double dt, double Re,
    const double* frame_vel, int refresh, const double* bc_vel, double cg_tol, int cg_maxit, int* cg_iterations)
{
    NEED_CTX("iblgf_gpu_ifherk_step_ib");
    if (!ctx->ib) return fail(IBLGF_ERR_STATE, "iblgf_gpu_ifherk_step_ib: no immersed-boundary system", ctx);
    if (!u_fields) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ifherk_step_ib: null velocity fields", ctx);
    if (!(dt > 0.0) || !(Re > 0.0)) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ifherk_step_ib: dt and Re must be positive", ctx);
    if (refresh < 0 || refresh > 2) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ifherk_step_ib: refresh must be 0, 1 or 2", ctx);
    if (!(cg_tol > 0.0) || cg_maxit < 1) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ifherk_step_ib: bad CG limits", ctx);
    return guarded(ctx, [&] {
        for (int c = 0; c < 3; ++c) ctx->solver->field_data(u_fields[c]);
        if (p_field >= 0) ctx->solver->field_data(p_field);
        ctx->ib->set_step_params(bc_vel, cg_tol, cg_maxit);
        ctx->solver->ifherk_step(u_fields, p_field, dt, Re, frame_vel, refresh, ctx->ib.get());
        if (cg_iterations)
        {
            ctx->solver->sync();
            for (int k = 0; k < 3; ++k) cg_iterations[k] = ctx->ib->stage_stats(k + 1).iterations;
        }
    });
}

int iblgf_gpu_ib_stage_forces(iblgf_gpu_ctx* ctx, double* forces, int write)
{
    NEED_CTX("iblgf_gpu_ib_stage_forces");
    if (!ctx->ib) return fail(IBLGF_ERR_STATE, "iblgf_gpu_ib_stage_forces: no immersed-boundary system", ctx);
    if (!forces) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ib_stage_forces: null array", ctx);
    return guarded(ctx, [&] {
        if (write) ctx->ib->write_stage_forces(forces);
        else
            ctx->ib->read_stage_forces(forces);
    });
}

int iblgf_gpu_ib_set_boundary_velocity(iblgf_gpu_ctx* ctx, const double* velocity)
{
    NEED_CTX("iblgf_gpu_ib_set_boundary_velocity");
    if (!ctx->ib) return fail(IBLGF_ERR_STATE, "iblgf_gpu_ib_set_boundary_velocity: no immersed-boundary system", ctx);
    return guarded(ctx, [&] { ctx->ib->set_point_velocities(velocity); });
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
            if (kernel == IBLGF_KERNEL_IF)
                return fail(IBLGF_ERR_UNSUPPORTED,
                    "iblgf_gpu_apply_lgf_typed: STREAM is built for the Laplace kernel only (the reference calls "
                    "apply_lgf with it, ifherk.hpp:1555)",
                    ctx);
            break;
        case IBLGF_FMM_IB2AMR:
            if (kernel == IBLGF_KERNEL_IF)
                return fail(IBLGF_ERR_UNSUPPORTED,
                    "iblgf_gpu_apply_lgf_typed: IB2AMR is built for the Laplace kernel only (the reference calls "
                    "apply_lgf with it, linsys.hpp:142)",
                    ctx);
            break;
        case IBLGF_FMM_LAPLACIAN_BC:
        case IBLGF_FMM_REF_FOURIER_STREAM:
        case IBLGF_FMM_REF_FOURIER_LGF:
            return fail(IBLGF_ERR_UNSUPPORTED,
                "iblgf_gpu_apply_lgf_typed: fmm_type " + std::to_string(fmm_type) +
                    " has no device path (built: AMR2AMR, IB2xIB, xIB2IB, STREAM, IB2AMR)",
                ctx);
        default:
            return fail(IBLGF_ERR_ARG, "iblgf_gpu_apply_lgf_typed: unknown fmm_type", ctx);
    }
    if (fmm_type == IBLGF_FMM_AMR2AMR) return iblgf_gpu_apply_lgf(ctx, src_field, dst_field, kernel, alpha_base);
    if (fmm_type == IBLGF_FMM_STREAM || fmm_type == IBLGF_FMM_IB2AMR)
    {
        NEED_CTX("iblgf_gpu_apply_lgf_typed");
        return guarded(ctx, [&] {
            ctx->solver->field_ptr(src_field);
            ctx->solver->field_ptr(dst_field);
            if (fmm_type == IBLGF_FMM_STREAM) ctx->solver->apply_lgf_masked(src_field, dst_field, MASK_STREAM);
            else
                ctx->solver->apply_lgf_ib2amr(src_field, dst_field);
        });
    }
    return iblgf_gpu_apply_lgf_masked(ctx, src_field, dst_field, kernel, alpha_base, fmm_type);
}

int iblgf_gpu_apply_lgf_noalloc(iblgf_gpu_ctx* ctx, int src_field, int dst_field, int kernel, double alpha_base)
{
    NEED_CTX("iblgf_gpu_apply_lgf_noalloc");
    if (kernel != IBLGF_KERNEL_LAPLACE && kernel != IBLGF_KERNEL_IF)
        return fail(IBLGF_ERR_ARG, "iblgf_gpu_apply_lgf_noalloc: unknown kernel", ctx);
    GpuSolver& s = *ctx->solver;
    const bool laplace = kernel == IBLGF_KERNEL_LAPLACE;
    const double alpha = laplace ? 0.0 : alpha_base;
    return guarded(ctx, [&] {
        const size_t need = s.region_bytes();
        if (need > kNoallocScratchBytes)
            throw capi::Unsupported("iblgf_gpu_apply_lgf_noalloc: the solve needs " + std::to_string(need) +
                                    " bytes of work memory; the fixed scratch holds " +
                                    std::to_string(kNoallocScratchBytes));
        if (!s.prepared(laplace, alpha))
            throw capi::StateError("iblgf_gpu_apply_lgf_noalloc: call iblgf_gpu_prepare(ctx, kernel, alpha_base) "
                                   "in the current precision first (plans and kernel spectra are allocated there)");
        void* region = noalloc_scratch(s.device());
        if (!region) throw capi::StateError("iblgf_gpu_apply_lgf_noalloc: no scratch address for this device");
        s.solve_in_region(src_field, dst_field, laplace, alpha, region, kNoallocScratchBytes);
    });
}

int iblgf_gpu_ib_create(iblgf_gpu_ctx* ctx, const iblgf_tree* t)
{
    NEED_CTX("iblgf_gpu_ib_create");
    if (!t || !t->has_ib) return fail(IBLGF_ERR_ARG, "iblgf_gpu_ib_create: tree without immersed-boundary points", ctx);
    if (ctx->solver->num_trees() != 1)
        return fail(IBLGF_ERR_UNSUPPORTED, "iblgf_gpu_ib_create: the context holds a forest (one tree only)", ctx);
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
