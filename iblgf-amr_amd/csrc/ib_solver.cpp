#include "ib_solver.hpp"

#include "ops.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <stdexcept>

namespace iblgf_amd {

#define IBLGF_HIP_IB(x)                                                                               \
    do                                                                                                \
    {                                                                                                 \
        const hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_)); \
    } while (0)

IbSystem::IbSystem(GpuSolver& solver, const IbGeometry& g) : s_(solver)
{
    if (solver.nblocks() == 0) throw std::runtime_error("tree not uploaded");
    npoints_ = g.size();
    if (npoints_ <= 0) throw std::invalid_argument("no immersed-boundary points");
    st_ = build_ib_stencil(solver.tree(), solver.slot_of(), g);
    d_sp_ptr_.upload(st_.spread_ptr);
    d_sp_dst_.upload(st_.spread_dst);
    d_sp_col_.upload(st_.spread_col);
    d_sp_w_.upload(st_.spread_w);
    d_in_ptr_.upload(st_.interp_ptr);
    d_in_node_.upload(st_.interp_node);
    d_in_w_.upload(st_.interp_w);
    std::vector<double> own(3 * static_cast<size_t>(npoints_));
    for (int i = 0; i < npoints_; ++i)
        for (int c = 0; c < 3; ++c) own[3 * i + c] = st_.owned[i] ? 1.0 : 0.0;
    d_owned_.upload(own);
    for (int c = 0; c < 3; ++c)
    {
        field_[c] = s_.field_create();
        aux_[c] = s_.field_create();
    }
    cell_ = s_.field_create();
    for (auto& v : d_tmp_) v.alloc(3 * static_cast<size_t>(npoints_));
    d_part_.alloc(launch::kDotBlocks);
    d_out_.alloc(1);
}

void IbSystem::smear(const double* d_force, const int fields[3])
{
    for (int c = 0; c < 3; ++c) s_.domain_clean(fields[c]);
    double* f[3] = {s_.field_ptr(fields[0]), s_.field_ptr(fields[1]), s_.field_ptr(fields[2])};
    launch::ib_spread(s_.stream(), d_force, d_sp_ptr_.p, d_sp_dst_.p, d_sp_col_.p, d_sp_w_.p,
        static_cast<int>(st_.spread_ptr.size()) - 1, f);
    IBLGF_HIP_IB(hipGetLastError());
}

void IbSystem::smear_add(const int fields[3])
{
    double* f[3] = {s_.field_ptr(fields[0]), s_.field_ptr(fields[1]), s_.field_ptr(fields[2])};
    launch::ib_spread(s_.stream(), d_force_.p, d_sp_ptr_.p, d_sp_dst_.p, d_sp_col_.p, d_sp_w_.p,
        static_cast<int>(st_.spread_ptr.size()) - 1, f, true);
    IBLGF_HIP_IB(hipGetLastError());
}

void IbSystem::init_stage_forces()
{
    if (d_fprev_.p) return;
    const size_t n3 = 3 * static_cast<size_t>(npoints_);
    d_fprev_.alloc(3 * n3);
    IBLGF_HIP_IB(hipMemset(d_fprev_.p, 0, 3 * n3 * sizeof(double)));
    d_force_.alloc(n3);
    d_uc_.alloc(n3);
    IBLGF_HIP_IB(hipMemset(d_force_.p, 0, n3 * sizeof(double)));
}

void IbSystem::read_stage_forces(double* host)
{
    init_stage_forces();
    IBLGF_HIP_IB(hipStreamSynchronize(s_.stream()));
    IBLGF_HIP_IB(hipMemcpy(host, d_fprev_.p, 9 * static_cast<size_t>(npoints_) * sizeof(double),
        hipMemcpyDeviceToHost));
}

void IbSystem::write_stage_forces(const double* host)
{
    init_stage_forces();
    IBLGF_HIP_IB(hipStreamSynchronize(s_.stream()));
    IBLGF_HIP_IB(hipMemcpy(d_fprev_.p, host, 9 * static_cast<size_t>(npoints_) * sizeof(double),
        hipMemcpyHostToDevice));
}

void IbSystem::set_step_params(const double* bc, double tol, int maxit)
{
    if (!(tol > 0.0) || maxit < 1) throw std::invalid_argument("CG tolerance and iteration limit must be positive");
    for (int i = 0; i < 9; ++i) bc_[i] = bc ? bc[i] : 0.0;
    step_tol_ = tol;
    step_maxit_ = maxit;
}

// LinSysSolver::subtract_boundary_vel (linsys.hpp:148-158) with bc_vel(idx,
// t, coordinate(i)) evaluated per point by the caller.
void IbSystem::set_point_velocities(const double* host)
{
    if (!host)
    {
        d_bcpts_.release();
        return;
    }
    std::vector<double> v(host, host + 9 * static_cast<size_t>(npoints_));
    IBLGF_HIP_IB(hipStreamSynchronize(s_.stream()));
    d_bcpts_.upload(v);
}

// LinSysSolver::ib_solve (linsys.hpp:113-129) inside the stage, with the
// stage's previous force as the initial guess (ifherk.hpp:1477-1484).
void IbSystem::force_solve(const int fields[3], double alpha, int stage, bool write_prev)
{
    if (stage < 1 || stage > 3) throw std::invalid_argument("stage out of range");
    const size_t n3 = 3 * static_cast<size_t>(npoints_);
    init_stage_forces();
    hipStream_t st = s_.stream();
    double*     prev = d_fprev_.p + (stage - 1) * n3;
    IBLGF_HIP_IB(hipMemcpyAsync(d_force_.p, prev, n3 * sizeof(double), hipMemcpyDeviceToDevice, st));
    project(fields, d_uc_.p);
    if (d_bcpts_.p)  // uc - v, as the reference subtracts bc_vel point by point
        launch::axpby(st, static_cast<long>(n3), -1.0, d_bcpts_.p + (stage - 1) * n3, 1.0, d_uc_.p, nullptr);
    else
        launch::force_shift(st, d_uc_.p, npoints_, bc_ + 3 * (stage - 1));
    for (int c = 0; c < 3; ++c) s_.domain_clean(fields[c]);
    stage_stats_[stage - 1] = cg(d_uc_.p, d_force_.p, alpha, step_tol_, step_maxit_);
    if (write_prev)
        IBLGF_HIP_IB(hipMemcpyAsync(prev, d_force_.p, n3 * sizeof(double), hipMemcpyDeviceToDevice, st));
}

void IbSystem::project(const int fields[3], double* d_force)
{
    const double* f[3] = {s_.field_ptr(fields[0]), s_.field_ptr(fields[1]), s_.field_ptr(fields[2])};
    launch::ib_interp(s_.stream(), f, d_in_ptr_.p, d_in_node_.p, d_in_w_.p, 3 * npoints_, d_force);
    IBLGF_HIP_IB(hipGetLastError());
}

// LinSysSolver::ET_H_S_E with Field = face_aux2 and fmm types IB2xIB
// (integrating factor) / xIB2IB (Schur complement), linsys.hpp:259-319.
void IbSystem::apply(const double* d_fin, double* d_fout, double alpha)
{
    const int L = s_.tree().depth - 1;
    for (int c = 0; c < 3; ++c)
    {
        s_.domain_clean(field_[c]);
        s_.domain_clean(aux_[c]);
    }
    smear(d_fin, field_);
    if (std::fabs(alpha) > 1e-14) s_.apply_if_masked_batch(field_, field_, 3, alpha, MASK_IB2xIB);
    s_.level_divergence(field_, cell_, L);
    s_.clean_ib_region_boundary(cell_, L, 2);
    s_.apply_lgf_masked(cell_, cell_, MASK_xIB2IB);
    s_.level_gradient(cell_, aux_, L);
    for (int c = 0; c < 3; ++c) s_.clean_ib_region_boundary(aux_[c], L, 2);
    for (int c = 0; c < 3; ++c) s_.add(aux_[c], field_[c], -1.0);
    project(field_, d_fout);
}

double IbSystem::dot(const double* a, const double* b)
{
    launch::dot(s_.stream(), a, b, d_owned_.p, 3L * npoints_, d_part_.p, d_out_.p);
    double v = 0.0;
    IBLGF_HIP_IB(hipMemcpyAsync(&v, d_out_.p, sizeof(double), hipMemcpyDeviceToHost, s_.stream()));
    IBLGF_HIP_IB(hipStreamSynchronize(s_.stream()));
    return v;
}

void IbSystem::add(double* a, const double* b, double s1, double s2)
{
    launch::axpby(s_.stream(), 3L * npoints_, s2, b, s1, a, d_owned_.p);
}

void IbSystem::set_recycling(int capacity)
{
    if (capacity < 0) throw std::invalid_argument("negative recycling capacity");
    rec_n_ = rec_next_ = 0;
    rec_cap_ = capacity;
    rec_dirty_ = false;
    if (capacity == 0)
    {
        rec_P_ = DevArray<double>();
        rec_Q_ = DevArray<double>();
        rec_qq_ = DevArray<double>();
        rec_coef_ = DevArray<double>();
        rec_qq_host_.clear();
        return;
    }
    const size_t len = 3 * static_cast<size_t>(npoints_);
    rec_P_.alloc(len * capacity);
    rec_Q_.alloc(len * capacity);
    rec_qq_.alloc(capacity);
    rec_coef_.alloc(capacity);
    rec_qq_host_.assign(capacity, 1.0);
}

void IbSystem::recycle_push(const double* p, const double* q, double qq)
{
    const size_t len = 3 * static_cast<size_t>(npoints_);
    const int    slot = rec_next_;
    IBLGF_HIP_IB(hipMemcpyAsync(rec_P_.p + slot * len, p, len * sizeof(double), hipMemcpyDeviceToDevice, s_.stream()));
    IBLGF_HIP_IB(hipMemcpyAsync(rec_Q_.p + slot * len, q, len * sizeof(double), hipMemcpyDeviceToDevice, s_.stream()));
    rec_qq_host_[slot] = qq;
    rec_dirty_ = true;
    rec_next_ = (rec_next_ + 1) % rec_cap_;
    rec_n_ = std::min(rec_n_ + 1, rec_cap_);
}

void IbSystem::recycle_sync_qq()
{
    if (!rec_dirty_) return;
    IBLGF_HIP_IB(hipMemcpy(rec_qq_.p, rec_qq_host_.data(), rec_cap_ * sizeof(double), hipMemcpyHostToDevice));
    rec_dirty_ = false;
}

// CG_solve (linsys.hpp:170-230).
CgStats IbSystem::cg(const double* d_uc, double* d_f, double alpha, double tol, int maxit)
{
    const auto t0 = std::chrono::steady_clock::now();
    CgStats    st;
    const long len = 3L * npoints_;
    double *   r = d_tmp_[0].p, *p = d_tmp_[1].p, *Ap = d_tmp_[2].p, *Ax = d_tmp_[3].p;
    apply(d_f, Ax, alpha);
    ++st.applies;
    IBLGF_HIP_IB(hipMemsetAsync(r, 0, len * sizeof(double), s_.stream()));
    add(r, d_uc, 0.0, 1.0);
    add(r, Ax, 1.0, -1.0);
    IBLGF_HIP_IB(hipMemcpyAsync(p, r, len * sizeof(double), hipMemcpyDeviceToDevice, s_.stream()));
    double rsold = dot(r, r);
    for (int k = 0; k < maxit; ++k)
    {
        apply(p, Ap, alpha);
        ++st.applies;
        const double pAp = dot(p, Ap);
        if (pAp == 0.0)
        {
            st.stop = "breakdown (p.Ap = 0)";
            break;
        }
        const double a = rsold / pAp;
        add(d_f, p, 1.0, a);
        add(r, Ap, 1.0, -a);
        const double rsnew = dot(r, r);
        const double f2 = dot(d_f, d_f);
        st.iterations = k + 1;
        st.residual = std::sqrt(rsnew / f2);
        st.history.push_back(st.residual);
        if (on_iteration) on_iteration(k, r, p, d_f);
        if (st.residual < tol)
        {
            st.stop = "converged";
            break;
        }
        add(p, r, rsnew / rsold, 1.0);
        rsold = rsnew;
    }
    if (st.stop.empty()) st.stop = "iteration limit";
    IBLGF_HIP_IB(hipStreamSynchronize(s_.stream()));
    st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return st;
}

// Flexible CG in its conjugate-residual form (flexible GCR; Eisenstat, Elman &
// Schultz 1983, with Notay's flexible preconditioning, SIAM J. Sci. Comput.
// 22 (2000) 1444-1460): z = M(r), q = A z, both orthogonalised against the
// kept directions in the q-inner product, alpha = (r.q)/(q.q).  Each step
// minimises |r| over the kept space, so the method converges for an operator
// whose symmetric part is positive definite even when it is only nearly
// symmetric, as the FMM-approximated force operator is.  kKeep directions are
// kept; when the store is full it restarts from the current residual.  A
// preconditioned direction that is not finite or not a descent direction
// (z.r <= 0) is replaced by r.
CgStats IbSystem::fcg(const double* d_uc, double* d_f, double alpha, double tol, int maxit, const Preconditioner& M,
    const FcgMonitor& mon)
{
    constexpr int kKeep = 30;
    const auto    t0 = std::chrono::steady_clock::now();
    CgStats       st;
    const long    len = 3L * npoints_;
    double *      r = d_tmp_[0].p, *Ax = d_tmp_[3].p;
    if ((int)krylov_p_.size() < kKeep)
    {
        krylov_p_.resize(kKeep);
        krylov_q_.resize(kKeep);
        for (auto& v : krylov_p_) v.alloc(len);
        for (auto& v : krylov_q_) v.alloc(len);
    }
    std::vector<double> qq;  // (q_i.q_i) of the kept directions
    const bool          recycle = rec_cap_ > 0;
    if (recycle && rec_n_ > 0 && rec_alpha_ != alpha) rec_n_ = rec_next_ = 0;  // another operator
    rec_alpha_ = alpha;
    apply(d_f, Ax, alpha);
    ++st.applies;
    IBLGF_HIP_IB(hipMemsetAsync(r, 0, len * sizeof(double), s_.stream()));
    add(r, d_uc, 0.0, 1.0);
    add(r, Ax, 1.0, -1.0);
    if (recycle && rec_n_ > 0)
    {
        // f += sum b_i p_i, r -= sum b_i q_i with b_i = (r.q_i) / (q_i.q_i)
        recycle_sync_qq();
        launch::multi_dot(s_.stream(), rec_Q_.p, rec_n_, len, r, d_owned_.p, rec_coef_.p);
        launch::multi_update(s_.stream(), d_f, 1.0, rec_P_.p, r, -1.0, rec_Q_.p, rec_coef_.p, rec_qq_.p, rec_n_, len,
            d_owned_.p);
    }
    // moves the current cycle's directions into the recycled store
    auto keep_cycle = [&]() {
        if (!recycle) return;
        for (size_t i = 0; i < qq.size(); ++i) recycle_push(krylov_p_[i].p, krylov_q_[i].p, qq[i]);
    };
    const bool monitored = mon.enabled && M && identity_rate_ > 0.0 && mon.window > 0;
    bool       use_m = static_cast<bool>(M);
    for (int k = 0; k < maxit; ++k)
    {
        if ((int)qq.size() == kKeep)  // restart: the cycle's store is full
        {
            keep_cycle();
            qq.clear();
        }
        const int j = static_cast<int>(qq.size());
        double*   p = krylov_p_[j].p;
        double*   q = krylov_q_[j].p;
        bool      identity = !use_m;
        if (use_m)
        {
            IBLGF_HIP_IB(hipMemsetAsync(p, 0, len * sizeof(double), s_.stream()));
            M(s_.stream(), r, p);
            const double zr = dot(p, r);
            if (!std::isfinite(zr) || zr <= 0.0)
            {
                identity = true;
                st.fell_back = true;
            }
        }
        if (identity) IBLGF_HIP_IB(hipMemcpyAsync(p, r, len * sizeof(double), hipMemcpyDeviceToDevice, s_.stream()));
        // non-owned entries of p are zeroed: they never enter the update
        launch::mask_zero(s_.stream(), len, d_owned_.p, p);
        apply(p, q, alpha);
        ++st.applies;
        if (recycle && rec_n_ > 0)
        {
            // orthogonalise against the recycled pairs: p -= sum b_i p_i, q -= sum b_i q_i
            recycle_sync_qq();
            launch::multi_dot(s_.stream(), rec_Q_.p, rec_n_, len, q, d_owned_.p, rec_coef_.p);
            launch::multi_update(s_.stream(), p, -1.0, rec_P_.p, q, -1.0, rec_Q_.p, rec_coef_.p, rec_qq_.p, rec_n_,
                len, d_owned_.p);
        }
        for (int i = 0; i < j; ++i)
        {
            const double beta = dot(q, krylov_q_[i].p) / qq[i];
            add(p, krylov_p_[i].p, 1.0, -beta);
            add(q, krylov_q_[i].p, 1.0, -beta);
        }
        const double qn = dot(q, q);
        if (!(qn > 0.0))
        {
            st.stop = "breakdown (q.q = 0)";
            break;
        }
        const double a = dot(r, q) / qn;
        add(d_f, p, 1.0, a);
        add(r, q, 1.0, -a);
        qq.push_back(qn);
        const double rr = dot(r, r), f2 = dot(d_f, d_f);
        st.iterations = k + 1;
        st.residual = std::sqrt(rr / f2);
        st.history.push_back(st.residual);
        if (on_iteration) on_iteration(k, r, p, d_f);
        if (st.residual < tol)
        {
            st.stop = "converged";
            break;
        }
        const int it = k + 1;
        if (monitored && use_m && it % mon.window == 0 && it >= mon.window)
        {
            const double h0 = st.history[it - mon.window], h1 = st.history[it - 1];
            const double rate = (std::log10(h0) - std::log10(h1)) / mon.window;
            if (!(rate >= mon.slack * identity_rate_))
            {
                use_m = false;
                st.fell_back = true;
                st.switched_at = it;
            }
        }
    }
    if (st.stop.empty()) st.stop = "iteration limit";
    keep_cycle();
    if (!M && st.history.size() >= 2)
    {
        const double r0 = st.history.front(), r1 = st.history.back();
        if (r0 > 0.0 && r1 > 0.0)
            identity_rate_ = (std::log10(r0) - std::log10(r1)) / static_cast<double>(st.history.size() - 1);
    }
    IBLGF_HIP_IB(hipStreamSynchronize(s_.stream()));
    st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return st;
}

}  // namespace iblgf_amd
