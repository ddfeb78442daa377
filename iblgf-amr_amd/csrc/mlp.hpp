// Kernels of the learned preconditioner of the immersed-boundary force CG:
// a per-point 3-layer MLP (tanh hidden layers) over the residuals of each
// point's K nearest surface neighbours expressed in the point's local frame.
// Weights are stored in BF16 (round to nearest even) with FP32 master copies
// for the optimiser; every product accumulates in FP32.
//
// No reference counterpart: the reference solves the force system with plain
// CG (include/iblgf/solver/linsys/linsys.hpp:170-230).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace iblgf_amd {
namespace launch {

// X[i][0:3K) = frame_i * r[nbr[i][j]] * inv_s (j = 0..K-1, j = 0 is the point
// itself), X[i][3K:6K) = rel[i] (neighbour offsets in the frame, constant).
void mlp_features(hipStream_t s, int npts, int K, const double* r, const int* nbr, const float* frame,
    const float* rel, float inv_s, float* X);
// T[i] = frame_i * e[i] * inv (training targets in the local frame).
void mlp_targets(hipStream_t s, int npts, const double* e, const float* frame, float inv, float* T);
// z[i] = scale * frame_i^T * Y[i] (FP64 output of the preconditioner).
void mlp_output(hipStream_t s, int npts, const float* Y, const float* frame, double scale, double* z);

// C[M][N] = epilogue( sum_k A(m,k) B(k,n) ) with
//   A(m,k) = ta ? A[k*lda + m] : A[m*lda + k]                       (FP32)
//   B(k,n) = tb ? B[n*ldb + k] : B[k*ldb + n], FP32 or BF16 (b_bf16)
// epilogue: + bias[n] (if bias), tanh (if act == 1), * (1 - P[m][n]^2) (if P).
void mlp_gemm(hipStream_t s, int M, int N, int K, const float* A, int lda, bool ta, const void* B, int ldb, bool tb,
    bool b_bf16, const float* bias, int act, const float* P, float* C);
// dY = 2 (Y - T) / count, partial sums of (Y - T)^2 per block (deterministic).
constexpr int kLossBlocks = 128;
void mlp_loss(hipStream_t s, long count, const float* Y, const float* T, float* dY, float* partials);
// column sums of dZ [M][N] -> db[N] (bias gradient)
void mlp_colsum(hipStream_t s, int M, int N, const float* dZ, float* db);
// Adam step on n parameters; refreshes the BF16 copy (if wq).
void mlp_adam(hipStream_t s, int n, float* w, const float* g, float* m, float* v, uint16_t* wq, float lr, float b1,
    float b2, float eps, int step);
void mlp_to_bf16(hipStream_t s, int n, const float* w, uint16_t* wq);
// rows gather: dst[r][:] = src[idx[r]][:] for r < nrows (row length len)
void mlp_gather_rows(hipStream_t s, int nrows, int len, const float* src, const int* idx, float* dst);

}  // namespace launch
}  // namespace iblgf_amd
