// Host-side interface of the binary64 refinement kernels (k_f64.hip).
#pragma once
#include "mf_common.hpp"

namespace mfk {

enum { TRI_NONE = 0, TRI_LOWER = 1, TRI_UPPER = 2 };

// C = alpha * A * B + beta * C on strided binary64 views (ls unused).
// a_tri: A(i, k) is zero for k > i (LOWER) / k < i (UPPER); b_tri: B(k, j) is
// zero for k < j (LOWER) / k > j (UPPER); those entries are never read.
// c_lower: only i >= j of C is written. Batch entry z offsets A, B, C by
// z * za / zb / zc elements. ksplit > 1 (batch 1 only) writes partial sums
// to P (ksplit * m * n doubles) and reduces them in ascending slice order.
struct DGemmArgs {
    int64_t m, n, k;
    DV A, B, C;
    double alpha, beta;
    int beta_zero;
    int a_tri, b_tri, c_lower;
    int64_t za, zb, zc;
    int ksplit;
    double *P;
};

hipError_t dgemm(hipStream_t s, DGemmArgs g, int64_t batch);
// Cholesky of the nb x nb (nb <= 64) diagonal block at (j0, j0) of F and the
// inverse of its factor into Inv at (j0, j0).
hipError_t dpotrf_diag(hipStream_t s, const DV &F, const DV &Inv, int64_t j0, int nb, Status *st);
hipError_t dcopy(hipStream_t s, int64_t m, int64_t n, const DV &dst, const DV &src, int mirror);
// out[i] = sum_j |A(i, j)|; part: drow_abs_chunks() * m doubles of scratch.
hipError_t drow_abs_sum(hipStream_t s, int64_t m, int64_t n, const DV &A, int uplo, double *out, double *part);
int drow_abs_chunks();
// Binary64 LU building blocks. dlu_leaf factors columns [c0, c0 + w) over
// rows [c0, m) with partial pivoting (w <= dlu_leaf_width(m - c0); 0 = too
// many rows), applies the leaf's row interchanges to the columns [a0, a0 +
// na) and [b0, b0 + nb) (na + nb + w <= 64) and solves the unit lower
// triangle of the leaf on the top w rows of the second range. dlu_rows
// applies a block's interchanges ipiv[j0 .. j0 + kn) (kn <= 64) to the
// columns left and right of it and solves L11 on the right columns.
int dlu_leaf_width(int64_t m);
hipError_t dlu_leaf(hipStream_t s, const DV &F, int64_t m, int64_t c0, int w, int64_t *ipiv, Status *st, int64_t a0, int na, int64_t b0,
                    int nb);
hipError_t dlu_rows(hipStream_t s, const DV &F, int64_t n, int64_t j0, int kn, const int64_t *ipiv);
hipError_t dgather_rows(hipStream_t s, int64_t m, int64_t ncols, const DV &dst, const DV &src, const int64_t *perm);

}  // namespace mfk
