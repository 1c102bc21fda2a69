// Host-side interface of the Ozaki-split INT8 MFMA products (k_ozaki.hip).
#pragma once
#include "mf_common.hpp"

namespace mfk {

// Digits per operand for an N-limb result (7 bits each, 34 for Float64x4).
int oz_digits(int N);
int oz_chunks();
// digit rows are padded to a multiple of this many bytes
int oz_row_align();
// e[i] = ilogb(max_k |hi(X(i, k))|) + 3 (or a zero-row sentinel); part:
// oz_chunks() * m doubles of scratch. uplo 0 general, 1 / 2: X symmetric
// stored in its lower / upper triangle.
hipError_t oz_exponents(hipStream_t s, int64_t m, int64_t K, const DV &X, int uplo, int *e, double *part);
// S balanced base-128 digits of X(i, k) 2^(7S - e_i): dig[s*plane + i*ld + k],
// ld a multiple of 64 >= K (padding digits zero). frag: the fragment layout
// read by oz_gemm_narrow instead (16-row blocks of 1 KB k steps; plane >=
// ceil(m / 16) * 16 * ld).
hipError_t oz_split(int N, hipStream_t s, int64_t m, int64_t K, const DV &X, int uplo, const int *e, int S, int8_t *dig, int64_t ld,
                    int64_t plane, bool frag = false);
// Rounds the binary64 columns D(:, j) in place to 7T-bit fixed point (scale
// 2^(e_j - 7T)) and writes their T digits dig[t*plane + j*ld + k].
hipError_t oz_round_fp64(hipStream_t s, int64_t K, int64_t ncols, const DV &D, const int *e, int T, int8_t *dig, int64_t ld,
                         int64_t plane);
// out[u][j][i] (int32, m x n column-major per u) = sum over digit pairs
// s + t = u of sum_k dA_s(i, k) dB_t(j, k), for u = 0..U. K = padded digit
// row length (<= 8192, multiple of oz_row_align()).
hipError_t oz_gemm(hipStream_t s, int64_t m, int64_t n, int64_t K, const int8_t *dA, int64_t lda, int64_t planeA, int SA,
                   const int8_t *dB, int64_t ldb, int64_t planeB, int SB, int U, int32_t *out);
// The same planes (U = SA + oz_narrow_digits() - 2) for A's digits in the
// fragment layout (K = the ld of the split, plane = planeA) and B with
// oz_narrow_digits() digits: the refinement's product A * D.
int oz_narrow_digits();
hipError_t oz_gemm_narrow(hipStream_t s, int64_t m, int64_t n, int64_t K, const int8_t *dA, int64_t planeA, int SA, const int8_t *dB,
                          int64_t ldb, int64_t planeB, int32_t *out);
// acc = sum_u planes[u] 2^(ea_i + eb_j - 7(u + 2)) in N limbs (u ascending);
// mode 0: C = alpha*acc (+ beta*C unless beta_zero); mode 1: C = C - acc.
hipError_t oz_combine(int N, hipStream_t s, int64_t m, int64_t n, int U, const int32_t *planes, const int *ea, const int *eb,
                      const DV &C, const double *alpha, const double *beta, int beta_zero, int mode);

}  // namespace mfk
