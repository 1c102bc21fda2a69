// Host-side interface of the factorization kernels (k_factor.hip).
#pragma once
#include "mf_common.hpp"

namespace mfk {

hipError_t chol_panel(int N, hipStream_t s, const DV &A, int64_t first, int64_t last, Status *st);
// Pivot of column k over rows k..m-1; the row interchange is applied to the
// panel columns [c0, c1) only (lu_laswp brings it to the other columns).
hipError_t lu_pivot_swap(int N, hipStream_t s, const DV &A, int64_t k, int64_t m, int64_t c0, int64_t c1, int64_t *ipiv, Status *st);
// ipiv[first .. first + count) applied in order to columns [a0, a0 + acount)
// and [b0, b0 + bcount).
hipError_t lu_laswp(int N, hipStream_t s, const DV &A, int64_t first, int64_t count, const int64_t *ipiv, int64_t a0, int64_t acount,
                    int64_t b0, int64_t bcount);
hipError_t lu_scale_update(int N, hipStream_t s, const DV &A, int64_t k, int64_t last, int64_t m, Status *st);
hipError_t apply_ipiv(int N, hipStream_t s, const DV &B, int64_t count, int64_t nrhs, const int64_t *ipiv);
hipError_t ldlt_unblocked(int N, hipStream_t s, const DV &A, int64_t n, const DV &dsub, int64_t *pivots, uint8_t *blocks,
                          const double *alpha, Status *st);
// spec != 0: a panel that extends past reqlast (a 2x2 pivot at its last
// column) records st->info = -3 (kernels launched after it return at once)
// and its last column in st->aux.
hipError_t ldlt_panel(int N, hipStream_t s, const DV &A, int64_t n, int64_t pfirst, int64_t reqlast, const DV &dsub, int64_t *pivots,
                      uint8_t *blocks, const double *alpha, double *keepbuf, int spec, Status *st);
// The lazy panel over `workgroups` (1..64) co-resident workgroups: the same
// bits as ldlt_panel. buf holds ldlt_panel_grid_doubles(N, n) doubles (the
// column / row-max values of four sets and the per-workgroup partials);
// weighted_ld > 0: at the end the panel's weighted rows W = L21 D overwrite
// buf as ldlt_weighted would write them with W's limb stride
// (n - plast - 1) * weighted_ld. st->info = -2 if a grid barrier timed out.
// trunc != 0: the panel stops at its first column that needs the row-max
// pass (kPanelTruncated, that column in st->aux) without touching that
// column's entries; ldlt_restore then returns its columns before it to the
// values they had at the launch.
size_t ldlt_panel_grid_doubles(int N, int64_t n);
hipError_t ldlt_panel_grid(int N, hipStream_t s, const DV &A, int64_t n, int64_t pfirst, int64_t reqlast, const DV &dsub, int64_t *pivots,
                           uint8_t *blocks, const double *alpha, double *buf, int spec, int trunc, int64_t weighted_ld, Status *st,
                           int workgroups);
hipError_t ldlt_restore(int N, hipStream_t s, const DV &A, int64_t n, int64_t pfirst, int64_t kstop, const double *buf);
hipError_t ldlt_weighted(int N, hipStream_t s, const DV &A, int64_t pfirst, int64_t plast, int64_t tfirst, int64_t tcount,
                         const DV &dsub, const uint8_t *blocks, const DV &W, const Status *st);
hipError_t mirror_lower(int N, hipStream_t s, const DV &A, int64_t n, const Status *st);
hipError_t ldlt_pivots(int N, hipStream_t s, const DV &B, int64_t n, int64_t nrhs, const int64_t *pivots, const uint8_t *blocks,
                       int forward, int *bad);
hipError_t ldlt_dsolve(int N, hipStream_t s, const DV &B, int64_t n, int64_t nrhs, const DV &F, const DV &dsub, const uint8_t *blocks,
                       Status *st);
hipError_t col_absmax_region(int N, hipStream_t s, const DV &A, int64_t m, int64_t ncols, int lower_only, const DV &out);
hipError_t absmax_fold(int N, hipStream_t s, const DV &colmax, int64_t count, double *out);
hipError_t iota1(hipStream_t s, int64_t *p, int64_t n);
hipError_t fill_u8(hipStream_t s, uint8_t *p, int64_t n, uint8_t v);

}  // namespace mfk
