// Host-side interface of the GEMM-family tile kernel (k_blas3.hip).
#pragma once
#include "mf_common.hpp"

namespace mfk {

enum TileMode { MODE_AB = 0, MODE_SUB = 1 };

struct TileArgs {
    int mode;          // MODE_AB or MODE_SUB
    int net;           // mf::NET_STD or mf::NET_FUSED3 (x3, MODE_AB only)
    int lower;         // write only elements with i >= j (absolute indices)
    int beta_zero;     // MODE_AB: C is not read
    int tile;          // -1 auto, 0: 16x16 tiles, 1: 32x32 tiles
    int64_t row_lo;    // first output row handled by this launch
    int64_t rows;      // number of output rows handled (from row_lo)
    int64_t n;         // output columns
    int64_t K;         // reduction length
    int64_t status_limit;  // 0: return on any recorded failure; > 0: only on
                           // failures at (one-based) columns <= status_limit
    int swap_product;      // MODE_SUB: terms b(k,j)*a(i,k) instead of a*b
};

// C(i, j) for i in [row_lo, row_lo + rows), j in [0, n) from a(i, k) = A(i, k)
// and b(k, j) = B(k, j), all strided views. alpha/beta are host limb arrays.
hipError_t tile_launch_n(int N, hipStream_t s, const TileArgs &t, const DV &A, const DV &B, const DV &C,
                         const double *alpha, const double *beta, const Status *st);

// C(i, j) = alpha * (sum over k of A(i, k) B(k, j), MAC chain from zero) +
// beta * C(i, j) for i < rows, j < ncols <= 8 (lower: i >= j only): the
// per-element operations of a MODE_AB NET_STD tile launch with beta read, in
// a layout for narrow outputs (32 x 8 per workgroup). With wait_on the
// kernel first waits (bounded) for *wait_on >= wait_value, a counter advanced
// by work enqueued before it on another stream; past the bound it writes NaN
// and *err = 1.
hipError_t cols_launch_n(int N, hipStream_t s, int64_t rows, int64_t ncols, int64_t K, bool lower, const DV &A, const DV &B, const DV &C,
                         const double *alpha, const double *beta, const Status *st, int64_t status_limit,
                         const unsigned *wait_on = nullptr, unsigned wait_value = 0, unsigned *err = nullptr);

// The output tile a launch over rows x n outputs with reduction length K
// takes (requested < 0: the automatic choice): its index 0..3 as the
// launcher resolves it, and its rows and columns per 256-thread workgroup.
int tile_geometry(int N, int64_t rows, int64_t n, int64_t K, bool lower, int requested, int *tile_rows, int *tile_cols);

}  // namespace mfk
