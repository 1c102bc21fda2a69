// Host-side interface of trmm!, mfdot and the packed-storage kernels (k_level3x.hip).
#pragma once
#include "mf_common.hpp"

namespace mfk {

// C = alpha * (left ? op(A) Bc : Bc op(A)) in MFLA's trmm! order; X / Y are
// the (coefficient or copy) operands of the triangular product, K its inner
// dimension, lower the effective triangle of op(A).
hipError_t trmm(int N, hipStream_t s, int left, int lower, int unit, int64_t m, int64_t n, int64_t K, const DV &X, const DV &Y,
                const DV &C, const double *alpha);
// out[0..N) = sum over the column-major linear index of x*y, ascending (device buffer).
hipError_t dot(int N, hipStream_t s, int64_t rows, int64_t cols, const DV &x, const DV &y, double *out);
// dense lower triangle -> packed (unpack = 0) or packed -> dense lower (unpack = 1).
hipError_t pack_lower(int N, hipStream_t s, int64_t n, const DV &dense, double *packed, int64_t pls, int unpack);

}  // namespace mfk
