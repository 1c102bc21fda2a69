// Host-side interface of the level-2 / elementwise kernels (k_blas2.hip).
#pragma once
#include "mf_common.hpp"

namespace mfk {

hipError_t status_reset(hipStream_t s, Status *st);
// op: 0 add 1 sub 2 mul 3 div 4 sqrt 5 z+x*y 6 fused-x3 z+x*y 7 inv 8 z-x*y 9 abs 10 max 11 min 12 lt
hipError_t arith(int N, hipStream_t s, int op, int64_t count, const double *x, const double *y, const double *z, double *out);
int host_arith(int N, int op, int64_t count, const double *x, const double *y, const double *z, double *out);
hipError_t scale(int N, hipStream_t s, int64_t m, int64_t n, const DV &B, const double *alpha, const Status *st);
hipError_t zero_fill(int N, hipStream_t s, int64_t m, int64_t n, const DV &B);
hipError_t copy(int N, hipStream_t s, int64_t m, int64_t n, const DV &dst, const DV &src);
hipError_t aos_to_planar(int N, hipStream_t s, int64_t m, int64_t n, const DV &dst, const double *src, int64_t src_ld);
hipError_t planar_to_aos(int N, hipStream_t s, int64_t m, int64_t n, double *dst, int64_t dst_ld, const DV &src);
hipError_t finite_check(int N, hipStream_t s, int64_t m, int64_t n, const DV &A, int region, Status *st);
hipError_t gemv(int N, hipStream_t s, int trans, int64_t m, int64_t n, const DV &A, const DV &x, const DV &y, const double *alpha,
                const double *beta, int beta_zero, const Status *st);
hipError_t symv(int N, hipStream_t s, int upper, int64_t n, int64_t nrhs, const DV &A, const DV &X, const DV &Y, const double *alpha,
                const double *beta, int beta_zero);
hipError_t residual_mixed(int NS, int NR, hipStream_t s, int uplo, int64_t m, int64_t n, int64_t nrhs, const DV &A, const DV &X,
                          const DV &B, const DV &R);
hipError_t add_mixed(int NS, int NR, hipStream_t s, int64_t m, int64_t n, const DV &X, const DV &D);
hipError_t row_abs_sum(int N, hipStream_t s, int uplo, int64_t m, int64_t n, const DV &A, const DV &out);
hipError_t col_absmax(int N, hipStream_t s, int64_t m, int64_t ncols, const DV &X, const DV &out, int *fin);

}  // namespace mfk
