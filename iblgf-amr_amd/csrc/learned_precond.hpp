// Learned preconditioner of the immersed-boundary force system: z = M(r)
// approximates A^-1 r point by point with a 3-layer MLP (hidden width H,
// tanh) over the residuals of each point's K nearest surface neighbours and
// their offsets, all in the point's local frame (tangent, tangent, outward
// normal).  M(c r) = c M(r): the input is scaled by the residual's RMS and
// the output by the same factor times a learned-once output scale.
//
// Training data are the (residual, correction) pairs a CG solve produces:
// for every recorded iterate k, (r_k, f* - f_k) with f* the converged force,
// i.e. exact (b, A^-1 b) pairs of the operator.  Training runs on the device
// with the kernels of mlp.hip (Adam, BF16 weights, FP32 accumulation).
//
// No reference counterpart: the reference solves with plain CG
// (include/iblgf/solver/linsys/linsys.hpp:170-230).
#pragma once

#include "ib_solver.hpp"

#include <cstdint>
#include <vector>

namespace iblgf_amd {

class LearnedPrecond
{
  public:
    LearnedPrecond(IbSystem& sys, const IbGeometry& g, int K = 16, int hidden = 128, uint64_t seed = 7);
    ~LearnedPrecond();
    LearnedPrecond(const LearnedPrecond&) = delete;
    LearnedPrecond& operator=(const LearnedPrecond&) = delete;

    // Record every stride-th CG iterate of the next solves of the system.
    void begin_collect(int stride);
    // Turn the recorded iterates into pairs against the final force d_f.
    void end_collect(const double* d_f);
    int  npairs() const { return static_cast<int>(pair_x_.size()); }
    void clear_pairs();

    // Adam on random row batches from the stored pairs; returns the mean
    // loss of the last 10 steps.
    double train(int steps, int batch_rows, float lr);
    // z = M(r) on device vectors of length 3N.
    void apply(hipStream_t s, const double* d_r, double* d_z);

    // Checking interface (host arrays): loss and gradients of a batch.
    // X [rows][6K], T [rows][3]; grads in parameter order.
    double loss_and_grad(const std::vector<float>& X, const std::vector<float>& T, int rows,
        std::vector<float>& grads);
    std::vector<float> params() const;     // FP32 master weights
    void               set_params(const std::vector<float>& p);
    std::vector<float> features(const double* d_r);  // [N][6K] for a device residual
    int    feature_dim() const { return 6 * K_; }
    int    hidden() const { return H_; }
    int    npoints() const { return N_; }
    double output_scale() const { return cscale_; }
    double train_seconds() const { return train_seconds_; }

  private:
    void   forward(int rows, const float* X);  // fills A1_, A2_, Y_
    double backward(int rows, const float* X, const float* T);  // fills grads_, returns loss
    double residual_scale(const double* d_r);

    IbSystem& sys_;
    int       N_, K_, H_, D_;
    long      nparam_ = 0;
    long      off_[6];  // W1 b1 W2 b2 W3 b3
    uint64_t  rng_;
    double    cscale_ = 0.0;
    double    train_seconds_ = 0.0;
    int       step_ = 0;

    DevArray<int>    d_nbr_;
    DevArray<float>  d_frame_, d_rel_;
    DevArray<float>  d_w_, d_g_, d_m_, d_v_;
    DevArray<uint16_t> d_wq_;
    DevArray<float>  d_X_, d_A1_, d_A2_, d_Y_, d_dY_, d_dZ2_, d_dZ1_, d_T_, d_part_;
    DevArray<int>    d_idx_;
    int              cap_rows_ = 0;
    DevArray<double> d_e_, d_tmp_;

    std::vector<DevArray<float>> pair_x_, pair_t_;
    std::vector<DevArray<double>> rec_r_, rec_f_;
    int  stride_ = 0;
    bool collecting_ = false;
    void ensure_rows(int rows);
};

}  // namespace iblgf_amd
