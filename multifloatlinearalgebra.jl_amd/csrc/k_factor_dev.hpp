// Workgroup-level device routines shared by the single-matrix factorization
// kernels (k_factor.hip) and the batched persistent kernel (k_batched.hip):
// reductions with MFLA's tie-break rules, the scaled 2x2 solve, symmetric
// swaps and the unblocked Bunch-Kaufman LDLT body. Both callers therefore run
// the same per-element arithmetic.
#pragma once
#include <climits>
#include "mf_common.hpp"

namespace mfk {

// ------------------------------------------------------------ utilities
template <int N> __device__ __forceinline__ void lds_put(double *s, const mf::V<N> &v) {
#pragma unroll
    for (int l = 0; l < N; ++l) s[l] = v.x[l];
}
template <int N> __device__ __forceinline__ mf::V<N> lds_get(const double *s) {
    mf::V<N> v;
#pragma unroll
    for (int l = 0; l < N; ++l) v.x[l] = s[l];
    return v;
}
template <int N> __device__ __forceinline__ void swap_el(const DV &A, int64_t i1, int64_t j1, int64_t i2, int64_t j2) {
    mf::V<N> t = ld<N>(A, i1, j1);
    stv<N>(A, i1, j1, ld<N>(A, i2, j2));
    stv<N>(A, i2, j2, t);
}

// Workgroup argmax of |value| with MFLA's first-strict-maximum rule: the
// largest value wins, ties go to the smallest index (lu.jl:93-100,
// ldlt.jl:192-200). Every thread passes its best (value, index); idx < 0 means
// "no candidate". The rule is a total order on (value, index), so a wave
// butterfly (shuffles) and one pass over the per-wave winners give the same
// winner as any other reduction order. blockDim.x is a multiple of 64.
template <int N>
__device__ void wg_argmax(mf::V<N> &val, int64_t &idx, double *sv_buf, int64_t *si_buf) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, nw = (int)(blockDim.x >> 6);
    auto better = [](const mf::V<N> &b, int64_t ib, const mf::V<N> &a, int64_t ia) {
        if (ib < 0) return false;
        if (ia < 0) return true;
        if (mf::gt(b, a)) return true;
        return mf::eq(a, b) && ib < ia;
    };
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        mf::V<N> b;
#pragma unroll
        for (int l = 0; l < N; ++l) b.x[l] = __shfl_xor(val.x[l], off, 64);
        const int64_t ib = __shfl_xor(idx, off, 64);
        if (better(b, ib, val, idx)) { val = b; idx = ib; }
    }
    if (lane == 0) { lds_put<N>(&sv_buf[wave * N], val); si_buf[wave] = idx; }
    __syncthreads();
    val = lds_get<N>(&sv_buf[0]);
    idx = si_buf[0];
    for (int w = 1; w < nw; ++w) {
        const mf::V<N> b = lds_get<N>(&sv_buf[w * N]);
        const int64_t ib = si_buf[w];
        if (better(b, ib, val, idx)) { val = b; idx = ib; }
    }
    __syncthreads();
}

// Workgroup max (Julia max semantics, value only).
template <int N>
__device__ mf::V<N> wg_max(mf::V<N> val, double *sv_buf) {
    const int t = threadIdx.x, T = blockDim.x;
    lds_put<N>(&sv_buf[t * N], val);
    __syncthreads();
    for (int s = T / 2; s > 0; s >>= 1) {
        if (t < s) lds_put<N>(&sv_buf[t * N], mf::vmax(lds_get<N>(&sv_buf[t * N]), lds_get<N>(&sv_buf[(t + s) * N])));
        __syncthreads();
    }
    mf::V<N> r = lds_get<N>(&sv_buf[0]);
    __syncthreads();
    return r;
}

__device__ int64_t wg_min_i64(int64_t v, int64_t *si_buf) {
    const int t = threadIdx.x, T = blockDim.x;
    si_buf[t] = v;
    __syncthreads();
    for (int s = T / 2; s > 0; s >>= 1) {
        if (t < s && si_buf[t + s] < si_buf[t]) si_buf[t] = si_buf[t + s];
        __syncthreads();
    }
    int64_t r = si_buf[0];
    __syncthreads();
    return r;
}

// wg_argmax with a quicker wave step for the usual case: when a wave's
// candidates have finite leading limbs and exactly one of them holds the
// largest, that lane is the wave's winner under the same rule (every other
// candidate is smaller in the leading limb); otherwise the full butterfly
// runs. A wave without candidates keeps idx < 0. Same result as wg_argmax.
template <int N>
__device__ void wg_argmax_lead(mf::V<N> &val, int64_t &idx, double *sv_buf, int64_t *si_buf) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, nw = (int)(blockDim.x >> 6);
    auto better = [](const mf::V<N> &b, int64_t ib, const mf::V<N> &a, int64_t ia) {
        if (ib < 0) return false;
        if (ia < 0) return true;
        if (mf::gt(b, a)) return true;
        return mf::eq(a, b) && ib < ia;
    };
    const bool cand = idx >= 0;
    const double h = val.x[0];
    const unsigned long long cands = __ballot(cand);
    if (cands != 0) {
        double m = cand ? h : -__builtin_inf();
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const double o = __shfl_xor(m, off, 64);
            m = o > m ? o : m;
        }
        const unsigned long long top = __ballot(cand && h == m);
        const bool nan = __ballot(cand && !(h == h)) != 0;
        if (!nan && __popcll(top) == 1) {
            const int w = __ffsll((long long)top) - 1;
#pragma unroll
            for (int l = 0; l < N; ++l) val.x[l] = __shfl(val.x[l], w, 64);
            idx = __shfl(idx, w, 64);
        } else {
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                mf::V<N> b;
#pragma unroll
                for (int l = 0; l < N; ++l) b.x[l] = __shfl_xor(val.x[l], off, 64);
                const int64_t ib = __shfl_xor(idx, off, 64);
                if (better(b, ib, val, idx)) { val = b; idx = ib; }
            }
        }
    }
    if (lane == 0) { lds_put<N>(&sv_buf[wave * N], val); si_buf[wave] = idx; }
    __syncthreads();
    val = lds_get<N>(&sv_buf[0]);
    idx = si_buf[0];
    for (int w = 1; w < nw; ++w) {
        const mf::V<N> b = lds_get<N>(&sv_buf[w * N]);
        const int64_t ib = si_buf[w];
        if (better(b, ib, val, idx)) { val = b; idx = ib; }
    }
    __syncthreads();
}

// ------------------------------------------------------------------ LDLT
// ldlt.jl:65-101 (scaled 2x2 elimination)
template <int N>
__device__ bool solve2x2(const mf::V<N> &d11, const mf::V<N> &d21, const mf::V<N> &d22, const mf::V<N> &f, const mf::V<N> &s,
                         mf::V<N> &x1, mf::V<N> &x2) {
    mf::V<N> scale = mf::vmax(mf::vmax(mf::vabs(d11), mf::vabs(d21)), mf::vabs(d22));
    x1 = mf::zero<N>(); x2 = mf::zero<N>();
    if (mf::is_zero(scale) || !mf::is_finite(scale)) return false;
    mf::V<N> a = mf::div(d11, scale), b = mf::div(d21, scale), c = mf::div(d22, scale);
    mf::V<N> r1 = mf::div(f, scale), r2 = mf::div(s, scale);
    if (mf::ge(mf::vabs(a), mf::vabs(b))) {
        if (mf::is_zero(a)) return false;
        mf::V<N> t = mf::div(b, a);
        mf::V<N> u = mf::msub(c, t, b);
        if (mf::is_zero(u) || !mf::is_finite(u)) return false;
        x2 = mf::div(mf::msub(r2, t, r1), u);
        x1 = mf::div(mf::msub(r1, b, x2), a);
        return true;
    }
    if (mf::is_zero(b)) return false;
    mf::V<N> t = mf::div(a, b);
    mf::V<N> u = mf::msub(b, t, c);
    if (mf::is_zero(u) || !mf::is_finite(u)) return false;
    x2 = mf::div(mf::msub(r1, t, r2), u);
    x1 = mf::div(mf::msub(r2, c, x2), b);
    return true;
}

// solve2x2 split into the part that depends on the pivot block only
// (prepared once per step) and the part of one right-hand side; together
// they perform solve2x2's operations in its order, so the bits are equal.
template <int N>
struct Prepared2x2 {
    bool ok, first;                     // first: |a| >= |b|
    mf::V<N> scale, a, b, c, t, u;
};
template <int N>
__device__ Prepared2x2<N> prepare2x2(const mf::V<N> &d11, const mf::V<N> &d21, const mf::V<N> &d22) {
    Prepared2x2<N> p;
    p.scale = mf::vmax(mf::vmax(mf::vabs(d11), mf::vabs(d21)), mf::vabs(d22));
    p.ok = false;
    p.first = true;
    if (mf::is_zero(p.scale) || !mf::is_finite(p.scale)) return p;
    p.a = mf::div(d11, p.scale);
    p.b = mf::div(d21, p.scale);
    p.c = mf::div(d22, p.scale);
    p.first = mf::ge(mf::vabs(p.a), mf::vabs(p.b));
    if (p.first) {
        if (mf::is_zero(p.a)) return p;
        p.t = mf::div(p.b, p.a);
        p.u = mf::msub(p.c, p.t, p.b);
    } else {
        if (mf::is_zero(p.b)) return p;
        p.t = mf::div(p.a, p.b);
        p.u = mf::msub(p.b, p.t, p.c);
    }
    p.ok = !(mf::is_zero(p.u) || !mf::is_finite(p.u));
    return p;
}
template <int N>
__device__ void apply2x2(const Prepared2x2<N> &p, const mf::V<N> &f, const mf::V<N> &s, mf::V<N> &x1, mf::V<N> &x2) {
    const mf::V<N> r1 = mf::div(f, p.scale), r2 = mf::div(s, p.scale);
    if (p.first) {
        x2 = mf::div(mf::msub(r2, p.t, r1), p.u);
        x1 = mf::div(mf::msub(r1, p.b, x2), p.a);
    } else {
        x2 = mf::div(mf::msub(r1, p.t, r2), p.u);
        x1 = mf::div(mf::msub(r2, p.c, x2), p.b);
    }
}

// ldlt.jl:155-181, executed cooperatively by one workgroup.
template <int N>
__device__ void wg_sym_swap(const DV &A, int64_t n, int64_t f, int64_t s, int64_t active) {
    if (f == s) return;
    for (int64_t c = threadIdx.x; c < n; c += blockDim.x) swap_el<N>(A, f, c, s, c);
    __syncthreads();
    for (int64_t r = active + threadIdx.x; r < n; r += blockDim.x) swap_el<N>(A, r, f, r, s);
    __syncthreads();
}

// Shared pieces of the Bunch-Kaufman pivot choice (ldlt.jl:183-219 and the
// lazy variant :451-508). The entry source E gives the current Schur entry
// E.entry(i1, i2) (called by every thread). E.prep(col) (called by every
// thread) readies E.centry(r, col) = entry(r, col) for r > col; E.prow(row)
// readies E.rentry(row, c) = entry(row, c) for c < row; E.keep(r, v) sees
// each column-pass value and E.keep_row(r, v) each centry(r, imax) of the
// row-max pass. With E::kFuseDiag, E.centry_diag(r, k, v, d) gives
// centry(r, k) and entry(k, k) together and E.diag_entry(k) gives entry(k, k)
// at the start of a step to any subset of threads. The lazy panel forms the
// factors those passes share once; the right-looking body reads A.
struct BKChoice { int bs; int64_t pivot; };

template <int N, typename E>
__device__ BKChoice wg_select_bk(int64_t n, int64_t k, const mf::V<N> &alpha, E &e, double *sv_buf, int64_t *si_buf) {
    mf::V<N> absakk;
    mf::V<N> best = mf::zero<N>();
    int64_t bi = -1;
    if constexpr (!E::kFuseDiag) {
        absakk = mf::vabs(e.entry(k, k));
        if (k == n - 1) return BKChoice{mf::is_zero(absakk) ? 0 : 1, k};
        e.prep(k);
        for (int64_t r = k + 1 + threadIdx.x; r < n; r += blockDim.x) {
            const mf::V<N> v = e.centry(r, k);
            e.keep(r, v);
            mf::V<N> c = mf::vabs(v);
            if (bi < 0 || mf::gt(c, best)) { best = c; bi = r; }
        }
    } else {
        // entry(k, k) formed beside the thread's first row (two independent
        // chains in one loop); a thread without rows forms it alone
        if (k == n - 1) {
            absakk = mf::vabs(e.entry(k, k));
            return BKChoice{mf::is_zero(absakk) ? 0 : 1, k};
        }
        e.prep(k);
        const int64_t r0 = k + 1 + threadIdx.x;
        if (r0 < n) {
            mf::V<N> v, dg;
            e.centry_diag(r0, k, v, dg);
            absakk = mf::vabs(dg);
            e.keep(r0, v);
            best = mf::vabs(v);
            bi = r0;
        } else {
            absakk = mf::vabs(e.diag_entry(k));
        }
        for (int64_t r = r0 + blockDim.x; r < n; r += blockDim.x) {
            const mf::V<N> v = e.centry(r, k);
            e.keep(r, v);
            mf::V<N> c = mf::vabs(v);
            if (mf::gt(c, best)) { best = c; bi = r; }
        }
    }
    wg_argmax<N>(best, bi, sv_buf, si_buf);
    const mf::V<N> colmax = best;
    const int64_t imax = bi;
    if (mf::eq(mf::vmax(absakk, colmax), mf::zero<N>())) return BKChoice{0, k};
    if (mf::ge(absakk, mf::mul(alpha, colmax))) return BKChoice{1, k};
    mf::V<N> rm = mf::zero<N>();
    e.prow(imax);
    for (int64_t c = k + threadIdx.x; c < imax; c += blockDim.x) rm = mf::vmax(rm, mf::vabs(e.rentry(imax, c)));
    e.prep(imax);
    for (int64_t r = imax + 1 + threadIdx.x; r < n; r += blockDim.x) {
        const mf::V<N> v = e.centry(r, imax);
        e.keep_row(r, v);
        rm = mf::vmax(rm, mf::vabs(v));
    }
    const mf::V<N> rowmax = wg_max<N>(rm, sv_buf);
    if (mf::ge(absakk, mf::mul(mf::mul(alpha, colmax), mf::div(colmax, rowmax)))) return BKChoice{1, k};
    if (mf::ge(mf::vabs(e.entry(imax, imax)), mf::mul(alpha, rowmax))) return BKChoice{1, imax};
    return BKChoice{2, imax};
}

// Entries of the right-looking factorization: the trailing block is updated
// in place, so every entry is the stored lower-triangle element.
template <int N>
struct StoredEntries {
    DV A;
    __device__ mf::V<N> entry(int64_t i, int64_t j) const { return ld<N>(A, i, j); }
    __device__ void prep(int64_t) const {}
    __device__ void keep(int64_t, const mf::V<N> &) const {}
    __device__ void keep_row(int64_t, const mf::V<N> &) const {}
    __device__ mf::V<N> centry(int64_t r, int64_t c) const { return ld<N>(A, r, c); }
    static constexpr bool kFuseDiag = false;
    __device__ void prow(int64_t) const {}
    __device__ mf::V<N> rentry(int64_t r, int64_t c) const { return ld<N>(A, r, c); }
};

// (row, col) = (e mod w, e div w), in 32-bit arithmetic when w * w fits.
__device__ __forceinline__ void sq_index(int64_t e, int64_t w, int64_t &row, int64_t &col) {
    if (w * w < 2147483647LL) {
        const unsigned ue = (unsigned)e, uw = (unsigned)w;
        col = ue / uw;
        row = ue - (unsigned)col * uw;
    } else {
        col = e / w;
        row = e - col * w;
    }
}

// Symmetric interchange f <-> s (f <= s) of ldlt.jl:155-181 on the lower
// triangle of a matrix whose block from `active` on is symmetric: the same
// lower-triangle values MFLA's full row + column swap leaves (swap rows f, s
// left of f; the transposed pairs (r, f) <-> (s, r) for f < r < s; columns
// f, s below s; the two diagonals).
template <int N>
__device__ void wg_sym_swap_lower(const DV &A, int64_t n, int64_t f, int64_t s) {
    if (f == s) return;
    for (int64_t c = threadIdx.x; c < f; c += blockDim.x) swap_el<N>(A, f, c, s, c);
    for (int64_t r = f + 1 + threadIdx.x; r < s; r += blockDim.x) swap_el<N>(A, r, f, s, r);
    for (int64_t r = s + 1 + threadIdx.x; r < n; r += blockDim.x) swap_el<N>(A, r, f, r, s);
    if (threadIdx.x == 0) swap_el<N>(A, f, f, s, s);
    __syncthreads();
}

// Upper triangle of the block [k0, n) from its lower triangle.
template <int N>
__device__ void wg_mirror_block(const DV &A, int64_t n, int64_t k0) {
    const int64_t w = n - k0;
    for (int64_t e = threadIdx.x; e < w * w; e += blockDim.x) {
        int64_t r, c;
        sq_index(e, w, r, c);
        if (r > c) stv<N>(A, k0 + c, k0 + r, ld<N>(A, k0 + r, k0 + c));
    }
    __syncthreads();
}

// Unblocked right-looking Bunch-Kaufman LDLT (ldlt.jl:352-412), one
// workgroup per matrix. Matrices are addressed through a batch stride so the
// same body serves single and batched factorizations. MFLA forms one
// coefficient per trailing column (coef = d * L(col, k), or c1 / c2 of the
// 2 x 2 step) before its row loop; with cbuf (2 * n * N doubles of LDS, or
// null) they are formed once per column here too, otherwise per element
// (same value either way).
//
// MFLA updates both triangles at every step. Its upper row i is never
// touched again once column i is finalized, and at that moment it equals
// the lower column i before scaling (the trailing block is symmetric after
// each interchange). So the updates here write the lower triangle only, each
// upper row is written once, when its column is finalized, and a failure
// mirrors the still-symmetric trailing block: every element of the returned
// matrix equals MFLA's, with O(n^2) instead of O(n^3) strided stores.
template <int N>
__device__ void ldlt_unblocked_body(const DV &A, int64_t n, const DV &dsub, int64_t *pivots, uint8_t *blocks, const mf::V<N> &alpha,
                                    int64_t *info_out, double *sv_buf, int64_t *si_buf, double *cbuf = nullptr) {
    StoredEntries<N> entries{A};
    auto fail = [&](int64_t code, int64_t k0) {
        wg_mirror_block<N>(A, n, k0);
        if (threadIdx.x == 0) *info_out = code;
    };
    int64_t k = 0;
    while (k < n) {
        BKChoice ch = wg_select_bk<N>(n, k, alpha, entries, sv_buf, si_buf);
        if (ch.bs == 0) { fail(k + 1, k); return; }
        if (ch.bs == 1) {
            if (threadIdx.x == 0) { pivots[k] = ch.pivot + 1; blocks[k] = 1; }
            wg_sym_swap_lower<N>(A, n, k, ch.pivot);
            const mf::V<N> d = ld<N>(A, k, k);
            if (mf::is_zero(d)) { fail(k + 1, k); return; }
            for (int64_t r = k + 1 + threadIdx.x; r < n; r += blockDim.x) {
                const mf::V<N> pre = ld<N>(A, r, k);
                stv<N>(A, k, r, pre);                                       // MFLA's frozen upper row k
                stv<N>(A, r, k, mf::div(pre, d));
            }
            __syncthreads();
            // rank-1 update (ldlt.jl:221-262), lower triangle
            const int64_t w = n - k - 1;
            if (cbuf) {
                for (int64_t j = threadIdx.x; j < w; j += blockDim.x) lds_put<N>(&cbuf[j * N], mf::mul(d, ld<N>(A, k + 1 + j, k)));
                __syncthreads();
            }
            for (int64_t e = threadIdx.x; e < w * w; e += blockDim.x) {
                int64_t row, col;
                sq_index(e, w, row, col);
                if (row < col) continue;
                const mf::V<N> coef = cbuf ? lds_get<N>(&cbuf[col * N]) : mf::mul(d, ld<N>(A, k + 1 + col, k));
                row += k + 1;
                col += k + 1;
                stv<N>(A, row, col, mf::msub(ld<N>(A, row, col), ld<N>(A, row, k), coef));
            }
            __syncthreads();
            k += 1;
        } else {
            if (!(k < n - 1)) { fail(k + 1, k); return; }
            if (threadIdx.x == 0) { pivots[k] = ch.pivot + 1; blocks[k] = 2; blocks[k + 1] = 0; }
            wg_sym_swap_lower<N>(A, n, k + 1, ch.pivot);
            const mf::V<N> d11 = ld<N>(A, k, k), d21 = ld<N>(A, k + 1, k), d22 = ld<N>(A, k + 1, k + 1);
            if (threadIdx.x == 0) stv<N>(dsub, k, 0, d21);
            mf::V<N> s1, s2;
            if (!solve2x2<N>(d11, d21, d22, mf::zero<N>(), mf::zero<N>(), s1, s2)) { fail(k + 1, k); return; }
            // MFLA's frozen upper rows k and k+1 (pre-solve lower columns)
            if (threadIdx.x == 0) stv<N>(A, k, k + 1, d21);
            for (int64_t r = k + 2 + threadIdx.x; r < n; r += blockDim.x) {
                stv<N>(A, k, r, ld<N>(A, r, k));
                stv<N>(A, k + 1, r, ld<N>(A, r, k + 1));
            }
            // rows k+2..n-1: find the first failing row, write only rows before it
            int64_t myfail = INT64_MAX;
            for (int64_t r = k + 2 + threadIdx.x; r < n; r += blockDim.x) {
                bool ok = solve2x2<N>(d11, d21, d22, ld<N>(A, r, k), ld<N>(A, r, k + 1), s1, s2);
                if (!(ok && mf::is_finite(s1) && mf::is_finite(s2)) && r < myfail) myfail = r;
            }
            const int64_t firstfail = wg_min_i64(myfail, si_buf);
            for (int64_t r = k + 2 + threadIdx.x; r < n && r < firstfail; r += blockDim.x) {
                solve2x2<N>(d11, d21, d22, ld<N>(A, r, k), ld<N>(A, r, k + 1), s1, s2);
                stv<N>(A, r, k, s1);
                stv<N>(A, r, k + 1, s2);
            }
            __syncthreads();
            if (firstfail != INT64_MAX) { fail(k + 1, k + 2); return; }
            // rank-2 update (ldlt.jl:264-319): vector form on four-row blocks,
            // scalar form on the remainder rows of each column; lower triangle
            if (k + 1 < n - 1) {
                const int64_t w = n - k - 2;
                auto coefs = [&](int64_t col, mf::V<N> &c1, mf::V<N> &c2) {
                    const mf::V<N> l1 = ld<N>(A, col, k), l2 = ld<N>(A, col, k + 1);
                    c1 = mf::add(mf::mul(d11, l1), mf::mul(d21, l2));
                    c2 = mf::add(mf::mul(d21, l1), mf::mul(d22, l2));
                };
                if (cbuf) {
                    for (int64_t j = threadIdx.x; j < w; j += blockDim.x) {
                        mf::V<N> c1, c2;
                        coefs(k + 2 + j, c1, c2);
                        lds_put<N>(&cbuf[j * N], c1);
                        lds_put<N>(&cbuf[(n + j) * N], c2);
                    }
                    __syncthreads();
                }
                for (int64_t e = threadIdx.x; e < w * w; e += blockDim.x) {
                    int64_t row, col;
                    sq_index(e, w, row, col);
                    if (row < col) continue;
                    mf::V<N> c1, c2;
                    if (cbuf) {
                        c1 = lds_get<N>(&cbuf[col * N]);
                        c2 = lds_get<N>(&cbuf[(n + col) * N]);
                    } else {
                        coefs(k + 2 + col, c1, c2);
                    }
                    row += k + 2;
                    col += k + 2;
                    const int64_t vec_end = col + ((n - col) / 4) * 4;
                    mf::V<N> v;
                    if (row < vec_end) v = mf::sub(ld<N>(A, row, col), mf::add(mf::mul(ld<N>(A, row, k), c1), mf::mul(ld<N>(A, row, k + 1), c2)));
                    else v = mf::msub(mf::msub(ld<N>(A, row, col), ld<N>(A, row, k), c1), ld<N>(A, row, k + 1), c2);
                    stv<N>(A, row, col, v);
                }
                __syncthreads();
            }
            if (threadIdx.x == 0) { stv<N>(A, k + 1, k, mf::zero<N>()); stv<N>(A, k, k + 1, mf::zero<N>()); }
            __syncthreads();
            k += 2;
        }
    }
}

}  // namespace mfk
