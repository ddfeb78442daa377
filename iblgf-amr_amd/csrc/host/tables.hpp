// Kernel and transfer tables of the LGF fast-multipole path (host, FP64).
//
//   Laplace LGF  G(n) = -int_0^inf prod_i e^{-2t} I_{n_i}(2t) dt for |n_i| <= N_max,
//                3-term asymptotic expansion beyond
//                (include/iblgf/lgf/lgf_gl_lookup.hpp:33-52, 74-119;
//                 include/iblgf/lgf/lgf_table_gen/lgf_integrator.py:7-25).
//   IF kernel    E_a(n) = e^{-6a} prod_i I_{n_i}(2a)  (include/iblgf/lgf/lgf_ge.hpp:158-168)
//   Nli          n x 2n barycentric Lagrange matrix, pts_cap = 10
//                (include/iblgf/fmm/fmm_nli.hpp:319-374)
//   cell-centre  Lagrange (relative position 0 / 1/2) and simple restriction
//                matrices (include/iblgf/interpolation/cell_center_nli_intrp.hpp:96-160, 677-787)
#pragma once

#include <vector>

namespace iblgf_amd {

// e^{-x} I_k(x) for k = 0..kmax (x >= 0).
void scaled_bessel_i(double x, int kmax, double* out);

struct LaplaceLgfTable
{
    int                 n_max = 100;
    std::vector<double> g;  // G for x>=y>=z>=0, index x(x+1)(x+2)/6 + y(y+1)/2 + z

    // Quadrature of the Bessel-product integral on t = e^s with step h,
    // s in [s_lo, ln t_hi], plus the two-term analytic tail beyond t_hi.
    void build(int n_max = 100, double h = 1.0 / 16.0, double s_lo = -42.0, double t_hi = 1e12);
    static double asym(int n1, int n2, int n3);
    double        get(int x, int y, int z) const;
    static long   index(int x, int y, int z)
    {
        return (long)x * (x + 1) * (x + 2) / 6 + (long)y * (y + 1) / 2 + z;
    }
};

// Integrating-factor kernel value e^{-6a} prod I_{n_i}(2a).
double if_kernel(int n1, int n2, int n3, double alpha);

// Nli transfer matrix: n rows (parent points) x 2n columns (child points).
// Columns [0,n) act on child 0, [n,2n) on child 1 along one axis.  Row-major
// storage mat[p*2n + c].
std::vector<double> nli_matrix(int n, int pts_cap = 10);

// cell_center_nli interpolation matrices, relative position 0 (coincident)
// and 1 (half-cell offset), n x 2n each, same column convention.
std::vector<double> cc_interp_matrix(int n, int pts_cap, int relative_pos);

// cell_center_nli restriction ("simple") sub-matrices, n x n, for child side
// c (0/1) and relative position rp (0: injection, 1: two-point average).
std::vector<double> cc_restrict_matrix(int n, int child_side, int relative_pos);

}  // namespace iblgf_amd
