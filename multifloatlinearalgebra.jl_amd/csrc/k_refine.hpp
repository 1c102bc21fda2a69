// Host-side interface of the refinement kernels (k_refine.hip).
#pragma once
#include "mf_common.hpp"

namespace mfk {

// Inverts the nb x nb (nb <= 64) diagonal blocks of the triangle of T
// (lower or upper, unit or not; binary64 only) into the same positions of W.
hipError_t trinv_blocks(int N, hipStream_t s, int64_t n, int nb, const DV &T, const DV &W, int lower, int unit);
// R (NR limbs) -= A (leading NA limbs) * D (binary64), products summed in NA
// limbs; uplo 0 general, 1 / 2 symmetric from the lower / upper triangle.
hipError_t resid_update(int NA, int NR, hipStream_t s, int uplo, int64_t m, int64_t nrhs, int64_t K, const DV &A, const DV &D, const DV &R);
// out[c] = max_i |X(i, c)| of a binary64 matrix.
hipError_t col_absmax_hi(hipStream_t s, int64_t m, int64_t ncols, const DV &X, double *out);

}  // namespace mfk
