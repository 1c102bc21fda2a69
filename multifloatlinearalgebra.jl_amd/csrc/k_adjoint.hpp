// Host-side interface of the adjoint element kernels (k_adjoint.hip).
#pragma once
#include "mf_common.hpp"

namespace mfk {

// out(i, j) = out(j, i) = alpha * (G(i, j) + G(j, i)); out may equal G.
hipError_t symmetrize(int N, hipStream_t s, int64_t n, const DV &G, const DV &out, const double *alpha);
// dst = lower triangle of src (diagonal halved if halve_diag), zero above.
hipError_t tril(int N, hipStream_t s, int64_t n, const DV &src, const DV &dst, int halve_diag);
// dst = identity
hipError_t identity(int N, hipStream_t s, int64_t n, const DV &dst);
// out[0..N) = m, out[N] = e with prod_i A(i, i) = m * 2^e, m in [1, 2) (device buffer).
hipError_t diag_logprod(int N, hipStream_t s, int64_t n, const DV &A, double *out);

}  // namespace mfk
