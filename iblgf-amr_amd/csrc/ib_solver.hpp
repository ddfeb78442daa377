// Immersed-boundary force system on the device: the operator
// A = E (I - G L^-1 D) H_alpha E^T of LinSysSolver::ET_H_S_E
// (include/iblgf/solver/linsys/linsys.hpp:259-292, apply_Schur 294-319) and
// its conjugate-gradient solve (CG_solve, 170-230), with FP64 dot products.
//
// Force vectors are N*3 FP64, point-major (the reference's force_type).
// Points not strictly inside a finest leaf block (ib rank -1 in the
// reference) are not owned: their residual is held at zero and they are
// left out of every dot product, as LinSysSolver::dot / add skip them.
#pragma once

#include "gpu_solver.hpp"
#include "host/ib.hpp"

#include <functional>
#include <string>
#include <vector>

namespace iblgf_amd {

struct CgStats
{
    int                 iterations = 0;
    double              residual = 0.0;  // sqrt(r.r / f.f) of the recurrence
    std::vector<double> history;         // residual after each iteration
    double              seconds = 0.0;
    int                 applies = 0;     // operator applications
    bool                fell_back = false;
    int                 switched_at = -1;  // iteration after which the identity replaced M
    std::string         stop;            // converged / iteration limit / breakdown
};

// Monitoring of a learned preconditioner inside the flexible solver: every
// `window` iterations the mean log10 residual reduction per iteration over the
// window is compared with `slack` times the rate of the last identity-
// preconditioned solve; below it, the identity replaces M for the rest of the
// solve (the reference's plain method is the fallback).
struct FcgMonitor
{
    bool   enabled = false;
    int    window = 10;
    double slack = 0.8;
};

// Flexible preconditioner: z = M(r) on device vectors of length 3N.
using Preconditioner = std::function<void(hipStream_t, const double* r, double* z)>;

class IbSystem : public StepIb
{
  public:
    IbSystem(GpuSolver& solver, const IbGeometry& g);

    int npoints() const { return npoints_; }
    hipStream_t stream() const { return s_.stream(); }
    const IbStencil& stencil() const { return st_; }
    const double*    owned_mask() const { return d_owned_.p; }

    // LinSysSolver::smearing / projection (linsys.hpp:350-376, 417-445)
    void smear(const double* d_force, const int fields[3]);
    void project(const int fields[3], double* d_force);
    // ET_H_S_E: fout = A fin
    void apply(const double* d_fin, double* d_fout, double alpha);

    // CG_solve: f holds the initial guess and receives the solution.
    CgStats cg(const double* d_uc, double* d_f, double alpha, double tol, int maxit);
    // Flexible CG in conjugate-residual form (flexible GCR, 30 kept
    // directions, restarted) with the preconditioner M (null: the identity);
    // a non-finite or non-descent preconditioned direction is replaced by r.
    // An identity solve records its convergence rate; a monitored solve
    // falls back to the identity when M converges slower than that rate.
    CgStats fcg(const double* d_uc, double* d_f, double alpha, double tol, int maxit, const Preconditioner& M,
        const FcgMonitor& mon = {});
    // Mean log10 residual reduction per iteration of the last identity solve
    // (0 before any).
    double identity_rate() const { return identity_rate_; }

    // Krylov recycling in fcg: the (p, A p) pairs of finished restart cycles
    // (mutually orthogonal in the owned inner product) are kept, up to
    // `capacity` pairs (oldest replaced first), while alpha stays the same.
    // Each solve first projects its residual onto the kept span, and every
    // new direction is orthogonalised against it, so later solves with the
    // same operator skip the slow modes earlier solves resolved.  0 turns it
    // off and drops the kept pairs.
    void set_recycling(int capacity);
    int  recycled() const { return rec_n_; }

    // FP64 dot over owned entries (LinSysSolver::dot).
    double dot(const double* a, const double* b);
    // a = a*s1 + b*s2 over owned entries (LinSysSolver::add).
    void add(double* a, const double* b, double s1, double s2);

    // The IF-HERK step's force solves (StepIb): the boundary velocity of
    // each stage (bc[9], x y z of stages 1..3, constant over the points), the
    // CG stopping rule, and the stage forces kept across steps (the
    // reference's ib().force_prev(stage), zero at first).
    void set_step_params(const double* bc, double tol, int maxit);
    // Per-point boundary velocities [3 stages][npoints][3] for the following
    // steps (null: back to the constant bc of set_step_params).
    void set_point_velocities(const double* host);
    void force_solve(const int fields[3], double alpha, int stage, bool write_prev) override;
    void smear_add(const int fields[3]) override;
    // the stage forces [3 stages][N][3] to or from the host
    void           read_stage_forces(double* host);
    void           write_stage_forces(const double* host);
    const CgStats& stage_stats(int stage) const { return stage_stats_[stage - 1]; }

    // Called after each CG/FCG iteration with (k, r, z or p, f) device
    // vectors; used to collect training pairs.
    std::function<void(int, const double*, const double*, const double*)> on_iteration;

  private:
    GpuSolver&      s_;
    int             npoints_ = 0;
    IbStencil       st_;
    DevArray<int>   d_sp_ptr_, d_sp_dst_, d_sp_col_, d_in_ptr_, d_in_node_;
    DevArray<double> d_sp_w_, d_in_w_, d_owned_;
    int             field_[3], aux_[3], cell_;
    DevArray<double> d_tmp_[6];  // r, p, Ap, Ax, z, spare
    DevArray<double> d_part_, d_out_;
    std::vector<DevArray<double>> krylov_p_, krylov_q_;
    double                        identity_rate_ = 0.0;

    void recycle_push(const double* p, const double* q, double qq);
    void recycle_sync_qq();
    DevArray<double>    d_fprev_, d_force_, d_uc_;
    DevArray<double>    d_bcpts_;  // per-point boundary velocities, empty when the constant bc_ holds
    void                init_stage_forces();
    double              bc_[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double              step_tol_ = 1e-3;
    int                 step_maxit_ = 100;
    CgStats             stage_stats_[3];
    DevArray<double>    rec_P_, rec_Q_, rec_qq_, rec_coef_;
    std::vector<double> rec_qq_host_;
    bool                rec_dirty_ = false;
    int                 rec_cap_ = 0, rec_n_ = 0, rec_next_ = 0;
    double              rec_alpha_ = 0.0;
};

}  // namespace iblgf_amd
