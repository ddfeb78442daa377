// IF-HERK time step on the device (Ifherk::time_step,
// include/iblgf/solver/time_integration/ifherk.hpp:1078-1200) for a tree
// without immersed-boundary points, built from the solver's batched LGF and
// integrating-factor solves and the block operators.
//
// Every operation of the step is a launch on the solver stream: the field
// operations stream whole blocks (HBM-bound), the stencils read the one-cell
// buffers that a halo fill refreshed, and the solves are the FMM path.  With
// graphs on, the step (about 400 launches on a 4000-block tree) is captured
// once per (fields, dt, Re, frame) and replayed as one graph.
#include "gpu_solver.hpp"
#include "kernels.hpp"
#include "ops.hpp"

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

namespace iblgf_amd {

#define IBLGF_HIP_HK(x)                                                                               \
    do                                                                                                \
    {                                                                                                 \
        const hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_)); \
    } while (0)

namespace {

// Operator::smooth2zero's ramp (operators.hpp:668-675).
double ramp(double x)
{
    const double fac = 10.0, shift = 0.2;
    const double c = 1 - (0.5 + 0.5 * std::tanh(fac * (1 - shift)));
    return ((0.5 + 0.5 * std::tanh(fac * (x - shift))) + c);
}

// The three-stage scheme of Ifherk (ifherk.hpp:2009-2010): a_ row by row, c_.
constexpr double kA[6] = {1.0 / 3, -1.0, 2.0, 0.0, 0.75, 0.25};
constexpr double kC[4] = {0.0, 1.0 / 3, 1.0, 1.0};
constexpr double coeff_a(int i, int j) { return kA[i * (i - 1) / 2 + j - 1]; }

}  // namespace

void GpuSolver::copy_scaled(int from, int to, double scale)
{
    launch::scale_copy(stream_, n_, F(to), F(from), p_all_.p, nb_, scale);
}

void GpuSolver::add_constant(int f, double v) { launch::add_const(stream_, n_, F(f), p_all_.p, nb_, v); }

void GpuSolver::step_lists()
{
    if (p_leaf_ring_.p) return;
    std::vector<int> nonleaf, ring;
    for (int s = 0; s < nb_; ++s)
    {
        if (flags_[s] & 1u) ring.insert(ring.end(), {s, 63});
        else
            nonleaf.push_back(s);
    }
    p_nonleaf_all_.upload(nonleaf);
    p_leaf_ring_.upload(ring);
    std::vector<double> r(2 * n_);
    for (int i = 0; i < n_; ++i)
    {
        const double pct = static_cast<double>(i) / static_cast<double>(n_ - 1);
        r[i] = ramp(pct);
        r[n_ + i] = ramp(1 - pct);
    }
    d_ramp_.upload(r);
}

void GpuSolver::clean_nonleaf(int f)
{
    step_lists();
    launch::zero_blocks(stream_, n_, F(f), p_nonleaf_all_.p, p_nonleaf_all_.size());
    launch::clean_faces(stream_, n_, F(f), p_leaf_ring_.p, p_leaf_ring_.size() / 2, 1);
}

const DevArray<int>& GpuSolver::smooth_list(bool leaf_only)
{
    const int key = leaf_only ? 1 : 0;
    auto      it = smooth_lists_.find(key);
    if (it != smooth_lists_.end()) return it->second;
    std::vector<int> ent;
    for (int s = lvl_begin_[base_level_]; s < lvl_begin_[base_level_ + 1]; ++s)
    {
        int mask = 0;
        for (int i = 0; i < 27; ++i)
        {
            const int nb = neighbor_[27 * s + i];
            if (nb < 0 || (leaf_only && (flags_[nb] & 2u))) mask |= 1 << i;
        }
        if (mask) ent.insert(ent.end(), {s, mask});
    }
    DevArray<int> a;
    a.upload(ent);
    return smooth_lists_.emplace(key, std::move(a)).first->second;
}

// Ifherk::clean_leaf_correction_boundary (ifherk.hpp:1352-1412) on one rank:
// with leaf_only the level's correction blocks are zeroed; at the base level
// every block is ramped towards each missing (or, with leaf_only, correction)
// neighbour.
void GpuSolver::smooth_boundary(int f, int level, bool leaf_only)
{
    if (leaf_only)
    {
        const auto& c = corr_list(level);
        launch::zero_blocks(stream_, n_, F(f), c.p, c.size());
    }
    if (level != base_level_) return;
    step_lists();
    const auto& l = smooth_list(leaf_only);
    launch::smooth_ramp(stream_, n_, F(f), l.p, l.size() / 2, d_ramp_.p);
}

// Ifherk::up (ifherk.hpp:999-1004) after clean<F>(true): every component
// coarsened (source_coarsify, poisson.hpp:1136-1194) with its mesh-type
// transfers.
void GpuSolver::step_up(const int* f, int k, int mesh)
{
    for (int c = 0; c < k; ++c) clean_nonleaf(f[c]);
    for (int c = 0; c < k; ++c)
    {
        Batch b;
        b.comp0 = c % 3;
        double* p = F(f[c]);
        for (int ls = depth_ - 2; ls >= base_level_; --ls)
            launch::transfer_gather(stream_, n_, p, p, p_coarsify_[ls].p, p_coarsify_[ls].size() / 9,
                m_cc_restrict_.p, true, b, false, mesh);
    }
}

// Ifherk::down_to_correction (ifherk.hpp:1006-1013): per component, each
// level's halo fill, then the interpolation into the correction children
// (intrp_to_correction_buffer, poisson.hpp:1015-1060).  Components are
// independent, so each runs its level loop on its own; but the reference's
// halo fill in that loop (buffer_exchange<From>) fills every component, so a
// component's buffers are filled once more, from its final values, during
// the loops of the components after it: here one more fill of each level.
void GpuSolver::step_down(const int* f, int k, int mesh)
{
    for (int c = 0; c < k; ++c)
    {
        Batch b;
        b.comp0 = c % 3;
        double* p = F(f[c]);
        for (int l = depth_ - 2; l >= base_level_; --l)
        {
            launch::buffer_exchange(stream_, n_, p, p_level_slots_[l].p, p_level_slots_[l].size(), d_neighbor_.p);
            launch::transfer_scatter(stream_, n_, p, p, p_ccint_corr_[l].p, p_ccint_corr_[l].size() / 3,
                m_cc_interp_.p, false, b, mesh);
        }
        if (c + 1 < k)
            for (int l = depth_ - 2; l >= base_level_; --l)
                launch::buffer_exchange(stream_, n_, p, p_level_slots_[l].p, p_level_slots_[l].size(), d_neighbor_.p);
    }
}

// Ifherk::up_and_down (ifherk.hpp:991-997)
void GpuSolver::up_and_down(const int* f, int k, int mesh)
{
    if (mesh != kMeshCell && mesh != kMeshFace && mesh != kMeshEdge) throw std::invalid_argument("unknown mesh type");
    step_up(f, k, mesh);
    step_down(f, k, mesh);
    IBLGF_HIP_HK(hipGetLastError());
}

const DevArray<int>& GpuSolver::noncorr_list(int level)
{
    auto it = noncorr_lists_.find(level);
    if (it != noncorr_lists_.end()) return it->second;
    std::vector<int> v;
    for (int s = lvl_begin_[level]; s < lvl_begin_[level + 1]; ++s)
        if (!(flags_[s] & 2u)) v.push_back(s);
    DevArray<int> a;
    a.upload(v);
    return noncorr_lists_.emplace(level, std::move(a)).first->second;
}

// Ifherk::pad_velocity<u, u>(true) (ifherk.hpp:1521-1575): the velocity
// coarsened; its curl on the non-correction blocks of every level, ramped
// at the base level; the base-level stream function from the STREAM solve
// of each component (fmm.hpp:199-233); the base-level velocity from its curl
// transpose times -1; then interpolated into the correction blocks.
void GpuSolver::step_pad_velocity(const int u[3])
{
    const StepFields& h = sf_;
    step_up(u, 3, kMeshFace);
    for (int c = 0; c < 3; ++c)
    {
        domain_clean(h.ea[c]);
        domain_clean(h.sf[c]);
    }
    for (int l = base_level_; l < depth_; ++l)
    {
        for (int c = 0; c < 3; ++c) buffer_exchange(u[c], l);
        const double  dx = dx_base_ / std::pow(2.0, l - base_level_);
        const double* sp[3] = {F(u[0]), F(u[1]), F(u[2])};
        double*       dp[3] = {F(h.ea[0]), F(h.ea[1]), F(h.ea[2])};
        const auto&   list = noncorr_list(l);
        launch::block_op(stream_, n_, OP_CURL, sp, dp, list.p, list.size(), 1.0 / dx);
    }
    for (int c = 0; c < 3; ++c) smooth_boundary(h.ea[c], base_level_, true);
    for (int c = 0; c < 3; ++c) masked_body(h.ea[c], h.sf[c], MASK_STREAM);
    block_op(OP_CURL_TRANSPOSE, h.sf, u, base_level_, -1.0);
    step_down(u, 3, kMeshFace);
}

// Ifherk::adapt_corr_time_step (ifherk.hpp:1206-1304): the three stages with
// alpha 0, without the nonlinear term.
void GpuSolver::step_adapt_corr(const int u[3])
{
    const StepFields& h = sf_;
    up_and_down(u, 3, kMeshFace);
    for (int c = 0; c < 3; ++c) copy_scaled(u[c], h.q[c], 1.0);
    for (int c = 0; c < 3; ++c)
    {
        domain_clean(h.g[c]);
        domain_clean(h.fa[c]);
    }
    domain_clean(h.di);
    domain_clean(h.ca);
    for (int c = 0; c < 3; ++c) copy_scaled(h.q[c], h.r[c], 1.0);
    step_lin_sys(0.0, 1, false);
    for (int c = 0; c < 3; ++c) domain_clean(h.r[c]);
    domain_clean(h.di);
    domain_clean(h.ca);
    for (int c = 0; c < 3; ++c) add(h.q[c], h.r[c], 1.0);
    up_and_down(h.ui, 3, kMeshFace);
    step_lin_sys(0.0, 2, false);
    domain_clean(h.di);
    domain_clean(h.ca);
    for (int c = 0; c < 3; ++c)
    {
        domain_clean(h.w2[c]);
        copy_scaled(h.q[c], h.r[c], 1.0);
    }
    mesh_ = kMeshFace;
    if_batch_body(h.r, h.r, 3, 0.0);
    mesh_ = kMeshCell;
    up_and_down(h.ui, 3, kMeshFace);
    step_lin_sys(0.0, 3, false);
    for (int c = 0; c < 3; ++c) copy_scaled(h.ui[c], u[c], 1.0);
}

void GpuSolver::nonlinear(const int u[3], const int w[3], const int d[3], int level, double scale)
{
    const double* up[3] = {F(u[0]), F(u[1]), F(u[2])};
    const double* wp[3] = {F(w[0]), F(w[1]), F(w[2])};
    double*       dp[3] = {F(d[0]), F(d[1]), F(d[2])};
    launch::nonlinear(stream_, n_, up, wp, dp, p_level_slots_[level].p, p_level_slots_[level].size(), scale);
}

// Ifherk::nonlinear (ifherk.hpp:1637-1703): vorticity by level, the velocity
// less the frame velocity, then the nonlinear term by level.
void GpuSolver::step_nonlinear(const int src[3], const int dst[3], double scale, const double frame[3])
{
    const StepFields& h = sf_;
    for (int c = 0; c < 3; ++c)
    {
        domain_clean(h.ea[c]);
        domain_clean(dst[c]);
        domain_clean(h.fa[c]);
    }
    for (int l = base_level_; l < depth_; ++l)
    {
        for (int c = 0; c < 3; ++c) buffer_exchange(src[c], l);
        block_op(OP_CURL, src, h.ea, l, 1.0);
    }
    for (int c = 0; c < 3; ++c)
    {
        copy_scaled(src[c], h.fa[c], 1.0);
        add_constant(h.fa[c], frame[c] * -1.0);
    }
    for (int l = base_level_; l < depth_; ++l)
    {
        for (int c = 0; c < 3; ++c) buffer_exchange(h.ea[c], l);
        for (int c = 0; c < 3; ++c) smooth_boundary(h.ea[c], l, false);
        nonlinear(h.fa, h.ea, dst, l, scale);
    }
}

// lin_sys_with_ib_solve (ifherk.hpp:1451-1517): divergence (1738-1764),
// Poisson solve, gradient (1766-1793); with IB points the IB2xIB
// integrating factor of r_i - grad d_i, the force solve, the pressure
// correction (linsys.hpp:131-145) and the gradient again with the smeared
// force added (without IB points these act on zero forces and drop out);
// then the subtraction and the integrating factor (or a copy).
void GpuSolver::step_lin_sys(double alpha, int stage, bool write_prev)
{
    const StepFields& h = sf_;
    up_and_down(h.r, 3, kMeshFace);
    for (int l = base_level_; l < depth_; ++l)
    {
        for (int c = 0; c < 3; ++c) buffer_exchange(h.r[c], l);
        block_op(OP_DIVERGENCE, h.r, &h.ca, l, 1.0);
        smooth_boundary(h.ca, l, true);
    }
    mesh_ = kMeshCell;
    lgf_batch_body(&h.ca, &h.di, 1);
    for (int c = 0; c < 3; ++c) domain_clean(h.fa[c]);
    for (int l = base_level_; l < depth_; ++l)
    {
        block_op(OP_GRADIENT, &h.di, h.fa, l, 1.0);
        for (int c = 0; c < 3; ++c) buffer_exchange(h.fa[c], l);
    }
    if (step_ib_)
    {
        for (int c = 0; c < 3; ++c)
        {
            copy_scaled(h.r[c], h.fa2[c], 1.0);
            add(h.fa[c], h.fa2[c], -1.0);
        }
        if (std::fabs(alpha) > 1e-14) apply_if_masked_batch(h.fa2, h.fa2, 3, alpha, MASK_IB2xIB);
        step_ib_->force_solve(h.fa2, alpha, stage, write_prev);
        // pressure_correction<d_i>
        for (int c = 0; c < 3; ++c) domain_clean(h.fa2[c]);
        domain_clean(h.ca2);
        step_ib_->smear_add(h.fa2);
        level_divergence(h.fa2, h.ca2, depth_ - 1);
        apply_lgf_ib2amr(h.ca2, h.ca2);
        add(h.ca2, h.di, -1.0);
        for (int c = 0; c < 3; ++c) domain_clean(h.fa[c]);
        for (int l = base_level_; l < depth_; ++l)
        {
            block_op(OP_GRADIENT, &h.di, h.fa, l, 1.0);
            for (int c = 0; c < 3; ++c) buffer_exchange(h.fa[c], l);
        }
        step_ib_->smear_add(h.fa);
    }
    for (int c = 0; c < 3; ++c) add(h.fa[c], h.r[c], -1.0);
    if (std::fabs(alpha) > 1e-14)
    {
        mesh_ = kMeshFace;
        if_batch_body(h.r, h.ui, 3, alpha);
        mesh_ = kMeshCell;
    }
    else
        for (int c = 0; c < 3; ++c) copy_scaled(h.r[c], h.ui[c], 1.0);
}

void GpuSolver::step_prepare(const double alpha[3], int refresh)
{
    if (!step_ready_)
    {
        StepFields h{};
        for (int c = 0; c < 3; ++c)
        {
            h.q[c] = field_create();
            h.r[c] = field_create();
            h.g[c] = field_create();
            h.w1[c] = field_create();
            h.w2[c] = field_create();
            h.ui[c] = field_create();
            h.fa[c] = field_create();
            h.ea[c] = field_create();
            h.sf[c] = field_create();
            h.fa2[c] = field_create();
        }
        h.di = field_create();
        h.ca = field_create();
        h.ca2 = field_create();
        sf_ = h;
        step_ready_ = true;
    }
    reserve_components(6);
    step_lists();
    smooth_list(false);
    smooth_list(true);
    for (int l = 0; l < depth_; ++l)
    {
        corr_list(l);
        noncorr_list(l);
    }
    prepare_laplace();
    prepare_if(alpha[0]);
    prepare_if(alpha[1]);
    if (std::fabs(alpha[2]) > 1e-14) prepare_if(alpha[2]);
    if (refresh)
    {
        // the STREAM pass and its kernel spectra
        PassPlan& sp = pass_idx(base_level_, fmm_mask_idx_gen(MASK_STREAM), false);
        for (auto& m : sp.mac) ghat(true, m.d, 0.0);
        if (refresh == 1) prepare_if(0.0);
        sync();
    }
}

void GpuSolver::step_body(const int u[3], int p, double dt, const double alpha[3], const double* frame, int refresh)
{
    const StepFields& h = sf_;
    // time_step (ifherk.hpp:1093-1116): the velocity refreshed from the
    // stream function (then, with use_adaptation_correction, the alpha-0
    // correction step), or up_and_down
    if (refresh)
    {
        step_pad_velocity(u);
        if (refresh == 1) step_adapt_corr(u);
    }
    else
        up_and_down(u, 3, kMeshFace);
    for (int c = 0; c < 3; ++c) copy_scaled(u[c], h.q[c], 1.0);

    // Stage 1
    for (int c = 0; c < 3; ++c)
    {
        domain_clean(h.g[c]);
        domain_clean(h.fa[c]);
    }
    domain_clean(h.di);
    domain_clean(h.ca);
    step_nonlinear(u, h.g, coeff_a(1, 1) * (-dt), frame);
    for (int c = 0; c < 3; ++c)
    {
        copy_scaled(h.q[c], h.r[c], 1.0);
        add(h.g[c], h.r[c], 1.0);
    }
    step_lin_sys(alpha[0], 1, true);

    // Stage 2
    for (int c = 0; c < 3; ++c) domain_clean(h.r[c]);
    domain_clean(h.di);
    domain_clean(h.ca);
    for (int c = 0; c < 3; ++c)
    {
        add(h.g[c], h.fa[c], -1.0);
        copy_scaled(h.fa[c], h.w1[c], -1.0 / dt / coeff_a(1, 1));
    }
    {
        // apply_lgf_IF<q_i, q_i> and apply_lgf_IF<w_1, w_1> at one alpha: one
        // batch of six face components, each solved in place
        const int qw[6] = {h.q[0], h.q[1], h.q[2], h.w1[0], h.w1[1], h.w1[2]};
        mesh_ = kMeshFace;
        if_batch_body(qw, qw, 6, alpha[0]);
        mesh_ = kMeshCell;
    }
    for (int c = 0; c < 3; ++c)
    {
        add(h.q[c], h.r[c], 1.0);
        add(h.w1[c], h.r[c], dt * coeff_a(2, 1));
    }
    up_and_down(h.ui, 3, kMeshFace);
    step_nonlinear(h.ui, h.g, coeff_a(2, 2) * (-dt), frame + 3);
    for (int c = 0; c < 3; ++c) add(h.g[c], h.r[c], 1.0);
    step_lin_sys(alpha[1], 2, true);

    // Stage 3
    domain_clean(h.di);
    domain_clean(h.ca);
    for (int c = 0; c < 3; ++c) domain_clean(h.w2[c]);
    for (int c = 0; c < 3; ++c)
    {
        add(h.g[c], h.fa[c], -1.0);
        copy_scaled(h.fa[c], h.w2[c], -1.0 / dt / coeff_a(2, 2));
        copy_scaled(h.q[c], h.r[c], 1.0);
        add(h.w1[c], h.r[c], dt * coeff_a(3, 1));
        add(h.w2[c], h.r[c], dt * coeff_a(3, 2));
    }
    mesh_ = kMeshFace;
    if_batch_body(h.r, h.r, 3, alpha[1]);
    mesh_ = kMeshCell;
    up_and_down(h.ui, 3, kMeshFace);
    step_nonlinear(h.ui, h.g, coeff_a(3, 3) * (-dt), frame + 6);
    for (int c = 0; c < 3; ++c) add(h.g[c], h.r[c], 1.0);
    step_lin_sys(alpha[2], 3, true);

    for (int c = 0; c < 3; ++c) copy_scaled(h.ui[c], u[c], 1.0);
    if (p >= 0) copy_scaled(h.di, p, 1.0 / coeff_a(3, 3) / dt);
    IBLGF_HIP_HK(hipGetLastError());
}

void GpuSolver::ifherk_step(const int u[3], int p, double dt, double Re, const double* frame, int refresh,
    StepIb* ib)
{
    if (refresh < 0 || refresh > 2) throw std::invalid_argument("refresh must be 0, 1 or 2");
    if (!uploaded_) throw std::runtime_error("tree not uploaded");
    if (ntrees_ != 1) throw std::invalid_argument("the time step runs on a single tree");
    if (!(dt > 0.0) || !(Re > 0.0)) throw std::invalid_argument("dt and Re must be positive");
    for (int c = 0; c < 3; ++c)
    {
        if (field_is_f32(u[c])) throw std::invalid_argument("the time step needs FP64 velocity fields");
        for (int e = 0; e < c; ++e)
            if (u[c] == u[e]) throw std::invalid_argument("velocity components must be distinct fields");
        if (p == u[c]) throw std::invalid_argument("the pressure field aliases a velocity component");
    }
    if (p >= 0 && field_is_f32(p)) throw std::invalid_argument("the time step needs an FP64 pressure field");
    IBLGF_HIP_HK(hipSetDevice(device_));
    double fr[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (frame) std::memcpy(fr, frame, sizeof(fr));
    // update_marching_parameters (ifherk.hpp:164-175)
    const double tmp = Re * dx_base_ * dx_base_ / dt;
    const double alpha[3] = {(kC[1] - kC[0]) / tmp, (kC[2] - kC[1]) / tmp, (kC[3] - kC[2]) / tmp};
    // the work fields exist before the key names them
    if (!step_ready_) step_prepare(alpha, refresh);
    std::vector<int64_t> key = {4, u[0], u[1], u[2], p, static_cast<int64_t>(precision_), refresh};
    auto bits = [](double x) {
        int64_t b;
        std::memcpy(&b, &x, sizeof(b));
        return b;
    };
    key.push_back(bits(dt));
    key.push_back(bits(Re));
    for (double x : fr) key.push_back(bits(x));
    try
    {
        if (ib)
        {
            // the force solves check convergence on the host: eager
            step_ib_ = ib;
            step_body(u, p, dt, alpha, fr, refresh);
            step_ib_ = nullptr;
        }
        else
            run_captured(key, [this, &alpha, refresh] { step_prepare(alpha, refresh); },
                [this, u, p, dt, &alpha, &fr, refresh] { step_body(u, p, dt, alpha, fr, refresh); });
    }
    catch (...)
    {
        step_ib_ = nullptr;
        mesh_ = kMeshCell;
        throw;
    }
    mesh_ = kMeshCell;
}

}  // namespace iblgf_amd
