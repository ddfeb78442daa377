// Launch interface of the block operators and immersed-boundary kernels.
// Fields use the layout of the LGF path: field[slot][n^3] FP64, x fastest,
// one array per vector component.  Every stencil writes the e^3 interior
// nodes (storage index 1..n-2 per axis) of the listed blocks and reads the
// one-cell buffer, so a halo fill precedes it where the reference does one.
//
// Arithmetic is not contracted into fused multiply-adds, so the results are
// the reference's operation sequence bit for bit.
#pragma once

#include <hip/hip_runtime.h>

namespace iblgf_amd {

enum BlockOp
{
    OP_LAPLACE = 0,         // Operator::laplace, operators.hpp:1387-1417         cell -> cell
    OP_GRADIENT = 1,        // Operator::gradient, operators.hpp:1425-1444        cell -> face
    OP_DIVERGENCE = 2,      // Operator::divergence, operators.hpp:1453-1474      face -> cell
    OP_CURL = 3,            // Operator::curl, operators.hpp:1483-1511            face -> edge
    OP_CURL_TRANSPOSE = 4,  // Operator::curl_transpose, operators.hpp:2420-2453  edge -> face
};

namespace launch {

// dst = op(src) over the interior of each listed block; fac = 1/dx (1/dx^2
// for the Laplacian, (1/dx)*scale for the curl transpose).
void block_op(hipStream_t s, int n, int op, const double* const src[3], double* const dst[3], const int* slots,
    int count, double fac);

// Zero `width` planes on the sides named by the 6-bit face mask of each
// entry {slot, mask}; bit b is neighbour slot {4,10,12,14,16,22}[b]
// (z-, y-, x-, x+, y+, z+).
void clean_faces(hipStream_t s, int n, double* f, const int* entries, int count, int width);
// Per listed block: the largest sum of squares of the k <= 3 components over
// its interior points (out[2b]) and the total of those sums (out[2b+1]).
void block_norms(hipStream_t s, int n, const double* const c[3], int k, const int* slots, int count, double* out);

// Operator::nonlinear (operators.hpp:2455-2500) on the listed blocks: d =
// 0.25 (w x u) on the interior nodes from face u and edge w (one-cell
// buffers read), then every node of d times scale.
void nonlinear(hipStream_t s, int n, const double* const u[3], const double* const w[3], double* const d[3],
    const int* slots, int count, double scale);
// dst[new slot] = src[old slot] for pairs {old, new}, whole blocks.
void remap_blocks(hipStream_t s, int n, double* dst, const double* src, const int* pairs, int count);
// dst = src * a and f = f + v on every node of the listed blocks.
void scale_copy(hipStream_t s, int n, double* dst, const double* src, const int* slots, int count, double a);
void add_const(hipStream_t s, int n, double* f, const int* slots, int count, double v);
// Operator::smooth2zero for entries {slot, 27-bit neighbour mask}; ramp[2][n]
// holds the ramp at node index idx towards the low and the high side.
void smooth_ramp(hipStream_t s, int n, double* f, const int* entries, int count, const double* ramp);

// Immersed-boundary transfer with precomputed delta weights.
// spread: field_c[node] = sum_k force[col_k] * w_k over the rows of node,
//         in stored order (rows list {comp, slot*n^3 + q}).
// interp: force[row] = sum_k field_c[node_k] * w_k, in stored order.
// add: the rows' sums start from the node's value (smearing without cleaning).
void ib_spread(hipStream_t s, const double* force, const int* row_ptr, const int* row_dst, const int* col,
    const double* w, int nrows, double* const fields[3], bool add = false);
// force[3 i + c] -= v[c] for every point (the boundary velocity subtracted).
void force_shift(hipStream_t s, double* force, int npoints, const double v[3]);
void ib_interp(hipStream_t s, const double* const fields[3], const int* row_ptr, const int* node, const double* w,
    int nrows, double* force);

// Deterministic FP64 dot product of two vectors of length len over the
// entries whose mask is nonzero (mask may be null): fixed two-stage
// reduction (partials[kDotBlocks] then one block), result in *out.
constexpr int kDotBlocks = 256;
void dot(hipStream_t s, const double* a, const double* b, const double* mask, long len, double* partials,
    double* out);
// y = y*b + x*a elementwise where mask is nonzero (mask may be null); the
// product and sum are rounded separately, as the reference's add().
void axpby(hipStream_t s, long len, double a, const double* x, double b, double* y, const double* mask);
// y = 0 where mask is zero.
void mask_zero(hipStream_t s, long len, const double* mask, double* y);
// out[i] = sum_j mask_j V[i][j] w[j] for nv rows of length len (deterministic).
void multi_dot(hipStream_t s, const double* V, int nv, long len, const double* w, const double* mask, double* out);
// x += sx * sum_i (c_i / d_i) X[i] and y += sy * sum_i (c_i / d_i) Y[i] where mask != 0;
// c and d are device arrays of nv values.
void multi_update(hipStream_t s, double* x, double sx, const double* X, double* y, double sy, const double* Y,
    const double* c, const double* d, int nv, long len, const double* mask);

}  // namespace launch
}  // namespace iblgf_amd
