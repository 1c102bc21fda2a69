// Host-side interface of the Householder QR kernels (k_qr.hip).
#pragma once
#include "mf_common.hpp"

namespace mfk {

constexpr int QR_PANEL = 16;   // qr.jl:303 (_QR_BLOCK_SIZE)

// Device state of one QR factorization. Vectors of Float64xN are stored
// limb-planar with limb stride `n` (scale, sum, tscale, tsum) or QR_PANEL (aux);
// F is the QR_PANEL x n delayed-update matrix of the blocked route (column
// stride QR_PANEL, limb stride QR_PANEL * n).
struct QrState {
    double *scale, *sum;     // column norm state (qr.jl:87-165)
    double *tscale, *tsum;   // speculative norm rebuilds of the hybrid pivot search
    int *dirty, *tflag;
    double *F, *aux, *diag;
    int64_t *pivot;          // selected pivot column (0-based)
    int *unreliable;         // any trailing dirty norm after the last update
};

// Norm state of columns [c0, c0 + ncols) over rows [first_row, m).
hipError_t qr_norms(int N, hipStream_t s, int64_t m, int64_t n, const DV &A, int64_t first_row, int64_t c0, int64_t ncols, const QrState &q);
// Householder reflector of column step (qr.jl:606-622) into A and tau(step);
// blocked = 1 additionally saves the diagonal into q.diag and writes 1 there.
hipError_t qr_reflector(int N, hipStream_t s, int64_t m, int64_t step, const DV &A, const DV &tau, int blocked, const QrState &q);
// T(:, c0 .. c0 + ncols) <- (I - tau(step) v v') T with v = (1, V(step+1 : m, vcol))
// (qr.jl:624-642, 890-929).
hipError_t qr_apply(int N, hipStream_t s, int64_t m, int64_t step, const DV &V, int64_t vcol, const DV &T, int64_t c0, int64_t ncols,
                    const DV &tau);
// Pivot search of the unblocked route (qr.jl:192-224) and of the blocked route
// (qr.jl:344-362); the result lands in *q.pivot.
hipError_t qr_pivot_hybrid(int N, hipStream_t s, int64_t m, int64_t n, int64_t step, const DV &A, const int64_t *perm, const QrState &q);
hipError_t qr_pivot_downdated(int N, hipStream_t s, int64_t n, int64_t step, const int64_t *perm, const QrState &q);
// Column interchange step <-> *q.pivot of A, of the first ls rows of F, of
// perm and of the norm state.
hipError_t qr_swap(int N, hipStream_t s, int64_t m, int64_t n, int64_t step, int64_t ls, const DV &A, int64_t *perm, const QrState &q);
// Blocked route, local step ls of the panel starting at bs (qr.jl:527-585):
// delayed prefix of the selected column, the auxiliary gemv, the new F row
// with the materialised R row.
hipError_t qr_delayed_column(int N, hipStream_t s, int64_t m, int64_t n, int64_t step, int64_t bs, int64_t ls, const DV &A, const QrState &q);
hipError_t qr_aux(int N, hipStream_t s, int64_t m, int64_t n, int64_t step, int64_t bs, int64_t ls, const DV &A, const DV &tau, const QrState &q);
hipError_t qr_frow(int N, hipStream_t s, int64_t m, int64_t n, int64_t step, int64_t bs, int64_t ls, const DV &A, const DV &tau, const QrState &q);
// Norm downdate after reflector step (qr.jl:264-301); pc > 0 rebuilds
// unreliable norms with the delayed prefix (F rows 0..pc-1 of the panel at
// bs). restore_diag writes q.diag back to A(step, step) first.
hipError_t qr_norm_update(int N, hipStream_t s, int64_t m, int64_t n, int64_t step, const DV &A, const DV &tau, int64_t bs, int64_t pc,
                          int restore_diag, const QrState &q);
// F = 0 (QR_PANEL x n), *q.unreliable = 0
hipError_t qr_panel_reset(int N, hipStream_t s, int64_t n, const QrState &q);

}  // namespace mfk
