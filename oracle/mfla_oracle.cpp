// TEST INFRASTRUCTURE ONLY -- the CPU oracle. A plain C++ restatement of the
// CPU algorithms of MultiFloatLinearAlgebra.jl (MFLA), with MFLA's per-element
// operation order. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it (through oracle/oracle.py), and only as the
// checker or as the timed CPU baseline; the product never links it.
//
// Storage is limb-planar: element (i, j), limb l of a matrix view lives at
// p[l * ls + j * ld + i] (column-major planes, zero-based indices). Pivot
// vectors and status codes follow MFLA's one-based convention.
//
// Each function names the reference file:line it restates. Parity pinning per
// function is listed in the header of oracle/oracle.py.
#include "mf_oracle.hpp"
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

using namespace orc;

namespace {

template <int N> struct View {
    double *p; int64_t ld, ls;
    inline MF<N> get(int64_t i, int64_t j) const { MF<N> r; for (int l = 0; l < N; ++l) r.x[l] = p[l * ls + j * ld + i]; return r; }
    inline void set(int64_t i, int64_t j, const MF<N> &v) const { for (int l = 0; l < N; ++l) p[l * ls + j * ld + i] = v.x[l]; }
    inline View sub(int64_t i, int64_t j) const { return View{p + j * ld + i, ld, ls}; }
};

template <int N> inline MF<N> scal(const double *s) { MF<N> r; for (int l = 0; l < N; ++l) r.x[l] = s[l]; return r; }

inline int nthreads_or(int t) {
#ifdef _OPENMP
    return t > 0 ? t : omp_get_max_threads();
#else
    (void)t; return 1;
#endif
}

// ----------------------------------------------------------------- GEMM
// src/kernels/gemm.jl:138-303 (direct) and :309-489 (fused x3): per element
// acc = 0; for k ascending acc = acc + A[i,k]*B[k,j]; then alpha*acc
// (beta == 0, C unread) or alpha*acc + beta*C. With the fused x3 route the
// four-row vector blocks use mulacc_x3 and the remainder rows (m mod 4) the
// standard network (gemm.jl:419-431, 444-451).
template <int N>
void gemm_t(int64_t m, int64_t n, int64_t k, View<N> A, View<N> B, View<N> C,
            MF<N> alpha, MF<N> beta, bool fused, int threads) {
    const bool overwrite = iszero(beta);
    const int64_t m4 = (m / 4) * 4;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads_or(threads))
    for (int64_t j = 0; j < n; ++j) {
        for (int64_t i = 0; i < m; ++i) {
            MF<N> acc = mf_zero<N>();
            if constexpr (N == 3) {
                if (fused && i < m4) {
                    for (int64_t kk = 0; kk < k; ++kk) acc = mulacc_x3(acc, A.get(i, kk), B.get(kk, j));
                } else {
                    for (int64_t kk = 0; kk < k; ++kk) acc = mac(acc, A.get(i, kk), B.get(kk, j));
                }
            } else {
                for (int64_t kk = 0; kk < k; ++kk) acc = mac(acc, A.get(i, kk), B.get(kk, j));
            }
            MF<N> r = overwrite ? mul(alpha, acc) : add(mul(alpha, acc), mul(beta, C.get(i, j)));
            C.set(i, j, r);
        }
    }
}

// ------------------------------------------------- SYRK / GEMMT (lower)
// src/kernels/syrk.jl:1-50 and src/kernels/gemmt.jl:1-52: for column c and
// rows r >= c, acc = sum_k left(r,k)*right(c,k) ascending (standard network on
// x86-64, syrk/gemmt `_structured_mulacc`, mulacc_x3.jl:98-116), then
// alpha*acc or alpha*acc + beta*C. Upper triangle untouched.
// left(r,k) = L[r,k] (ltrans=0) or L[k,r] (ltrans=1); same for right.
template <int N>
void lower_update_t(int64_t n, int64_t k, View<N> L, int ltrans, View<N> R, int rtrans, View<N> C,
                    MF<N> alpha, MF<N> beta, int threads) {
    const bool overwrite = iszero(beta);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads_or(threads))
    for (int64_t c = 0; c < n; ++c) {
        for (int64_t r = c; r < n; ++r) {
            MF<N> acc = mf_zero<N>();
            for (int64_t kk = 0; kk < k; ++kk) {
                MF<N> a = ltrans ? L.get(kk, r) : L.get(r, kk);
                MF<N> b = rtrans ? R.get(kk, c) : R.get(c, kk);
                acc = mac(acc, a, b);
            }
            MF<N> res = overwrite ? mul(alpha, acc) : add(mul(alpha, acc), mul(beta, C.get(r, c)));
            C.set(r, c, res);
        }
    }
}

// ----------------------------------------------------------------- TRSM
// src/kernels/trsm.jl:1-302. op(A) X = alpha B (left) or X op(A) = alpha B
// (right), in place. B is scaled by alpha first unless alpha == 1
// (trsm.jl:10-16). Each solution element: v = B; for k in substitution order
// v = v - coeff*B[k]; v = v / diag (unless unit).
template <int N>
void trsm_t(int side_right, int upper, int trans, int unit, int64_t m, int64_t n,
            View<N> A, View<N> B, MF<N> alpha, int threads) {
    if (!is_one(alpha)) {
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < m; ++i) B.set(i, j, mul(B.get(i, j), alpha));
    }
    const bool transposed = trans != 0;
    const bool eff_lower = ((!upper) == (!transposed));
    auto coeff = [&](int64_t row, int64_t col) { return transposed ? A.get(col, row) : A.get(row, col); };
    if (!side_right) {
        const int64_t na = m;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads_or(threads))
        for (int64_t j = 0; j < n; ++j) {
            if (eff_lower) {
                for (int64_t row = 0; row < na; ++row) {
                    MF<N> v = B.get(row, j);
                    for (int64_t kk = 0; kk < row; ++kk) v = msub(v, coeff(row, kk), B.get(kk, j));
                    if (!unit) v = mf_div(v, A.get(row, row));
                    B.set(row, j, v);
                }
            } else {
                for (int64_t row = na - 1; row >= 0; --row) {
                    MF<N> v = B.get(row, j);
                    for (int64_t kk = row + 1; kk < na; ++kk) v = msub(v, coeff(row, kk), B.get(kk, j));
                    if (!unit) v = mf_div(v, A.get(row, row));
                    B.set(row, j, v);
                }
            }
        }
    } else {
        const int64_t na = n;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads_or(threads))
        for (int64_t i = 0; i < m; ++i) {
            if (eff_lower) {
                for (int64_t col = na - 1; col >= 0; --col) {
                    MF<N> v = B.get(i, col);
                    for (int64_t kk = col + 1; kk < na; ++kk) v = msub(v, B.get(i, kk), coeff(kk, col));
                    if (!unit) v = mf_div(v, A.get(col, col));
                    B.set(i, col, v);
                }
            } else {
                for (int64_t col = 0; col < na; ++col) {
                    MF<N> v = B.get(i, col);
                    for (int64_t kk = 0; kk < col; ++kk) v = msub(v, B.get(i, kk), coeff(kk, col));
                    if (!unit) v = mf_div(v, A.get(col, col));
                    B.set(i, col, v);
                }
            }
        }
    }
}

// ----------------------------------------------------------------- GEMV
// src/kernels/gemv.jl:1-93: y = alpha*op(A)*x + beta*y, ascending reduction,
// beta == 0 => y unread.
// ------------------------------------------------------------------ TRMM
// src/kernels/trmm.jl:1-73. Per element acc = 0 and, with "+=" meaning
// acc = acc + x*y: left lower k = 0..i-1 then the diagonal term; left upper
// the diagonal term then k = i+1..m-1; right lower the diagonal then
// k = j+1..n-1; right upper k = 0..j-1 then the diagonal. Terms are coeff*B
// (left) or B*coeff (right); a unit diagonal adds B(i, j). Result alpha*acc.
// The in-place orders (rows descending for left lower, columns descending for
// right upper) keep every consumed input unmodified.
template <int N>
void trmm_t(int side_right, int upper, int trans, int unit, int64_t m, int64_t n, View<N> A, View<N> B, MF<N> alpha) {
    const bool transposed = trans != 0;
    const bool eff_lower = ((!upper) == (!transposed));
    auto coeff = [&](int64_t row, int64_t col) { return transposed ? A.get(col, row) : A.get(row, col); };
    if (!side_right) {
        for (int64_t j = 0; j < n; ++j)
            for (int64_t t = 0; t < m; ++t) {
                const int64_t i = eff_lower ? m - 1 - t : t;
                MF<N> acc = mf_zero<N>();
                auto diag = [&]() { acc = unit ? add(acc, B.get(i, j)) : add(acc, mul(A.get(i, i), B.get(i, j))); };
                if (eff_lower) {
                    for (int64_t k = 0; k < i; ++k) acc = add(acc, mul(coeff(i, k), B.get(k, j)));
                    diag();
                } else {
                    diag();
                    for (int64_t k = i + 1; k < m; ++k) acc = add(acc, mul(coeff(i, k), B.get(k, j)));
                }
                B.set(i, j, mul(alpha, acc));
            }
    } else {
        for (int64_t i = 0; i < m; ++i)
            for (int64_t t = 0; t < n; ++t) {
                const int64_t j = eff_lower ? t : n - 1 - t;
                MF<N> acc = mf_zero<N>();
                auto diag = [&]() { acc = unit ? add(acc, B.get(i, j)) : add(acc, mul(B.get(i, j), A.get(j, j))); };
                if (eff_lower) {
                    diag();
                    for (int64_t k = j + 1; k < n; ++k) acc = add(acc, mul(B.get(i, k), coeff(k, j)));
                } else {
                    for (int64_t k = 0; k < j; ++k) acc = add(acc, mul(B.get(i, k), coeff(k, j)));
                    diag();
                }
                B.set(i, j, mul(alpha, acc));
            }
    }
}

// ------------------------------------------------------------------ dot
// src/kernels/dot.jl:8-22: acc = acc + x[i]*y[i] over the column-major
// linear index, ascending.
template <int N> MF<N> dot_t(int64_t m, int64_t n, View<N> x, View<N> y) {
    MF<N> acc = mf_zero<N>();
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) acc = add(acc, mul(x.get(i, j), y.get(i, j)));
    return acc;
}

// ----------------------------------------------------------- packed SYRK
// src/kernels/syrk.jl:110-254: packed column-major lower triangle,
// (r, c) -> c(2n - c + 1)/2 + (r - c) (0-based); per element acc = 0,
// acc = acc + P[k, r]*P[k, c] over k = rf..rl ascending, then alpha*acc
// (beta == 0, output unread) or alpha*acc + beta*out.
template <int N>
void syrk_packed_t(int64_t n, int64_t rf, int64_t rl, View<N> P, double *out, int64_t ls, MF<N> alpha, MF<N> beta) {
    const bool bz = iszero(beta);
    for (int64_t c = 0; c < n; ++c)
        for (int64_t r = c; r < n; ++r) {
            MF<N> acc = mf_zero<N>();
            for (int64_t k = rf; k <= rl; ++k) acc = add(acc, mul(P.get(k, r), P.get(k, c)));
            const int64_t idx = c * (2 * n - c + 1) / 2 + (r - c);
            MF<N> o;
            for (int l = 0; l < N; ++l) o.x[l] = out[l * ls + idx];
            const MF<N> v = bz ? mul(alpha, acc) : add(mul(alpha, acc), mul(beta, o));
            for (int l = 0; l < N; ++l) out[l * ls + idx] = v.x[l];
        }
}

template <int N>
void gemv_t(int trans, int64_t m, int64_t n, View<N> A, View<N> x, View<N> y, MF<N> alpha, MF<N> beta) {
    const bool overwrite = iszero(beta);
    const int64_t outs = trans ? n : m, red = trans ? m : n;
    for (int64_t o = 0; o < outs; ++o) {
        MF<N> acc = mf_zero<N>();
        for (int64_t kk = 0; kk < red; ++kk) {
            MF<N> a = trans ? A.get(kk, o) : A.get(o, kk);
            acc = mac(acc, a, x.get(kk, 0));
        }
        y.set(o, 0, overwrite ? mul(alpha, acc) : add(mul(alpha, acc), mul(beta, y.get(o, 0))));
    }
}

// src/kernels/symv.jl:1-36: each row sums columns 1..row-1, the diagonal,
// then row+1..n, reading only the authoritative triangle.
template <int N>
void symv_t(int upper, int64_t n, View<N> A, View<N> x, View<N> y, MF<N> alpha, MF<N> beta) {
    const bool overwrite = iszero(beta);
    for (int64_t row = 0; row < n; ++row) {
        MF<N> acc = mf_zero<N>();
        for (int64_t col = 0; col < row; ++col) acc = mac(acc, upper ? A.get(col, row) : A.get(row, col), x.get(col, 0));
        acc = mac(acc, A.get(row, row), x.get(row, 0));
        for (int64_t col = row + 1; col < n; ++col) acc = mac(acc, upper ? A.get(row, col) : A.get(col, row), x.get(col, 0));
        y.set(row, 0, overwrite ? mul(alpha, acc) : add(mul(alpha, acc), mul(beta, y.get(row, 0))));
    }
}

// ------------------------------------------------------------- Cholesky
// src/factorizations/cholesky.jl:10-107.
template <int N> bool lower_finite(int64_t n, View<N> A) {
    for (int64_t c = 0; c < n; ++c)
        for (int64_t r = c; r < n; ++r)
            if (!isfin(A.get(r, c))) return false;
    return true;
}

template <int N>
int64_t potrf_t(int64_t n, View<N> A, int64_t block, int threads) {
    if (!lower_finite(n, A)) return -1;
    block = std::max<int64_t>(block, 1);
    const MF<N> one = mf_from<N>(1.0), mone = mf_from<N>(-1.0);
    for (int64_t first = 0; first < n; first += block) {
        const int64_t last = std::min(first + block, n) - 1;
        for (int64_t col = first; col <= last; ++col) {
            MF<N> d = A.get(col, col);
            for (int64_t kk = first; kk < col; ++kk) { MF<N> v = A.get(col, kk); d = msub(d, v, v); }
            if (!gt(d, mf_zero<N>())) return col + 1;
            MF<N> piv = mf_sqrt(d);
            A.set(col, col, piv);
            for (int64_t row = col + 1; row <= last; ++row) {
                MF<N> v = A.get(row, col);
                for (int64_t kk = first; kk < col; ++kk) v = msub(v, A.get(row, kk), A.get(col, kk));
                A.set(row, col, mf_div(v, piv));
            }
        }
        if (last + 1 < n) {
            const int64_t nb = last - first + 1, rest = n - last - 1;
            View<N> L11 = A.sub(first, first), A21 = A.sub(last + 1, first), A22 = A.sub(last + 1, last + 1);
            trsm_t<N>(1, 0, 1, 0, rest, nb, L11, A21, one, threads);
            lower_update_t<N>(rest, nb, A21, 0, A21, 0, A22, mone, one, threads);
        }
    }
    return 0;
}

// ------------------------------------------------------------------- LU
// src/factorizations/lu.jl:34-195.
template <int N>
int64_t getrf_t(int64_t m, int64_t n, View<N> A, int64_t block, int64_t *ipiv, double *origmax, int gemm_fused, int threads) {
    const int64_t kmax = std::min(m, n);
    for (int64_t i = 0; i < kmax; ++i) ipiv[i] = i + 1;
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i)
            if (!isfin(A.get(i, j))) { for (int l = 0; l < N; ++l) origmax[l] = 0.0; return -1; }
    MF<N> mx = mf_zero<N>();
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) mx = mf_max(mx, mf_abs(A.get(i, j)));
    for (int l = 0; l < N; ++l) origmax[l] = mx.x[l];
    block = std::max<int64_t>(block, 1);
    const MF<N> one = mf_from<N>(1.0), mone = mf_from<N>(-1.0);
    for (int64_t first = 0; first < kmax; first += block) {
        const int64_t last = std::min(first + block, kmax) - 1;
        for (int64_t k = first; k <= last; ++k) {
            int64_t piv = k;
            MF<N> pv = mf_abs(A.get(k, k));
            for (int64_t row = k + 1; row < m; ++row) {
                MF<N> cand = mf_abs(A.get(row, k));
                if (gt(cand, pv)) { piv = row; pv = cand; }
            }
            ipiv[k] = piv + 1;
            if (piv != k)
                for (int64_t c = 0; c < n; ++c) { MF<N> t = A.get(k, c); A.set(k, c, A.get(piv, c)); A.set(piv, c, t); }
            if (iszero(A.get(k, k))) return k + 1;
            MF<N> inv = mf_inv(A.get(k, k));
            for (int64_t row = k + 1; row < m; ++row) A.set(row, k, mul(A.get(row, k), inv));
            if (k < last) {
                for (int64_t c = k + 1; c <= last; ++c) {
                    MF<N> mult = A.get(k, c);
                    for (int64_t row = k + 1; row < m; ++row) A.set(row, c, msub(A.get(row, c), A.get(row, k), mult));
                }
            }
        }
        if (last + 1 < n) {
            const int64_t nb = last - first + 1;
            View<N> L11 = A.sub(first, first), U12 = A.sub(first, last + 1);
            trsm_t<N>(0, 0, 0, 1, nb, n - last - 1, L11, U12, one, threads);
            if (last + 1 < m) {
                View<N> A21 = A.sub(last + 1, first), A22 = A.sub(last + 1, last + 1);
                gemm_t<N>(m - last - 1, n - last - 1, nb, A21, U12, A22, mone, one, gemm_fused != 0, threads);
            }
        }
    }
    return 0;
}

// ----------------------------------------------------------------- LDLT
// src/factorizations/ldlt.jl:65-101
template <int N>
bool ldlt_solve_2x2(const MF<N> &d11, const MF<N> &d21, const MF<N> &d22, const MF<N> &first, const MF<N> &second,
                    MF<N> &x1, MF<N> &x2) {
    MF<N> scale = mf_max(mf_max(mf_abs(d11), mf_abs(d21)), mf_abs(d22));
    x1 = mf_zero<N>(); x2 = mf_zero<N>();
    if (iszero(scale) || !isfin(scale)) return false;
    MF<N> a = mf_div(d11, scale), b = mf_div(d21, scale), c = mf_div(d22, scale);
    MF<N> r1 = mf_div(first, scale), r2 = mf_div(second, scale);
    if (ge(mf_abs(a), mf_abs(b))) {
        if (iszero(a)) return false;
        MF<N> t = mf_div(b, a);
        MF<N> u = msub(c, t, b);
        if (iszero(u) || !isfin(u)) return false;
        x2 = mf_div(msub(r2, t, r1), u);
        x1 = mf_div(msub(r1, b, x2), a);
        return true;
    }
    if (iszero(b)) return false;
    MF<N> t = mf_div(a, b);
    MF<N> u = msub(b, t, c);
    if (iszero(u) || !isfin(u)) return false;
    x2 = mf_div(msub(r1, t, r2), u);
    x1 = mf_div(msub(r2, c, x2), b);
    return true;
}

template <int N> void swap_el(View<N> A, int64_t i1, int64_t j1, int64_t i2, int64_t j2) {
    MF<N> t = A.get(i1, j1); A.set(i1, j1, A.get(i2, j2)); A.set(i2, j2, t);
}

// ldlt.jl:155-181
template <int N> void ldlt_sym_swap(int64_t n, View<N> A, int64_t f, int64_t s, int64_t active) {
    if (f == s) return;
    for (int64_t c = 0; c < active; ++c) swap_el(A, f, c, s, c);
    for (int64_t c = active; c < n; ++c) swap_el(A, f, c, s, c);
    for (int64_t r = active; r < n; ++r) swap_el(A, r, f, r, s);
}

// ldlt.jl:183-219 (returns block size, sets pivot zero-based)
template <int N> int select_bk(int64_t n, View<N> A, int64_t k, const MF<N> &alpha, int64_t &pivot) {
    MF<N> absakk = mf_abs(A.get(k, k));
    pivot = k;
    if (k == n - 1) return iszero(absakk) ? 0 : 1;
    int64_t imax = k + 1;
    MF<N> colmax = mf_abs(A.get(imax, k));
    for (int64_t row = k + 2; row < n; ++row) {
        MF<N> cand = mf_abs(A.get(row, k));
        if (gt(cand, colmax)) { colmax = cand; imax = row; }
    }
    if (eq(mf_max(absakk, colmax), mf_zero<N>())) return 0;
    if (ge(absakk, mul(alpha, colmax))) return 1;
    MF<N> rowmax = mf_zero<N>();
    for (int64_t c = k; c < imax; ++c) rowmax = mf_max(rowmax, mf_abs(A.get(imax, c)));
    for (int64_t r = imax + 1; r < n; ++r) rowmax = mf_max(rowmax, mf_abs(A.get(r, imax)));
    if (ge(absakk, mul(mul(alpha, colmax), mf_div(colmax, rowmax)))) return 1;
    if (ge(mf_abs(A.get(imax, imax)), mul(alpha, rowmax))) { pivot = imax; return 1; }
    pivot = imax;
    return 2;
}

// ldlt.jl:352-412 (+ rank-1 :221-262, rank-2 :264-319)
template <int N>
int64_t ldlt_unblocked_t(int64_t n, View<N> A, View<N> dsub, int64_t *pivots, uint8_t *blocks, const MF<N> &alpha) {
    int64_t k = 0;
    while (k < n) {
        int64_t pivot;
        int bs = select_bk(n, A, k, alpha, pivot);
        if (bs == 0) return k + 1;
        if (bs == 1) {
            pivots[k] = pivot + 1; blocks[k] = 1;
            ldlt_sym_swap(n, A, k, pivot, k);
            MF<N> d = A.get(k, k);
            if (iszero(d)) return k + 1;
            for (int64_t row = k + 1; row < n; ++row) A.set(row, k, mf_div(A.get(row, k), d));
            if (k < n - 1) {
                for (int64_t col = k + 1; col < n; ++col) {
                    MF<N> coef = mul(d, A.get(col, k));
                    for (int64_t row = col; row < n; ++row) {
                        MF<N> v = msub(A.get(row, col), A.get(row, k), coef);
                        A.set(row, col, v); A.set(col, row, v);
                    }
                }
            }
            k += 1;
        } else {
            if (!(k < n - 1)) return k + 1;
            pivots[k] = pivot + 1; blocks[k] = 2; blocks[k + 1] = 0;
            ldlt_sym_swap(n, A, k + 1, pivot, k);
            MF<N> d11 = A.get(k, k), d21 = A.get(k + 1, k), d22 = A.get(k + 1, k + 1);
            dsub.set(k, 0, d21);
            MF<N> s1, s2;
            if (!ldlt_solve_2x2(d11, d21, d22, mf_zero<N>(), mf_zero<N>(), s1, s2)) return k + 1;
            for (int64_t row = k + 2; row < n; ++row) {
                bool ok = ldlt_solve_2x2(d11, d21, d22, A.get(row, k), A.get(row, k + 1), s1, s2);
                if (!(ok && isfin(s1) && isfin(s2))) return k + 1;
                A.set(row, k, s1); A.set(row, k + 1, s2);
            }
            if (k + 1 < n - 1) {
                for (int64_t col = k + 2; col < n; ++col) {
                    MF<N> l1 = A.get(col, k), l2 = A.get(col, k + 1);
                    MF<N> c1 = add(mul(d11, l1), mul(d21, l2));
                    MF<N> c2 = add(mul(d21, l1), mul(d22, l2));
                    const int64_t vec_end = col + ((n - col) / 4) * 4;  // four-row vector blocks
                    for (int64_t row = col; row < n; ++row) {
                        MF<N> v;
                        if (row < vec_end) v = sub(A.get(row, col), add(mul(A.get(row, k), c1), mul(A.get(row, k + 1), c2)));
                        else v = msub(msub(A.get(row, col), A.get(row, k), c1), A.get(row, k + 1), c2);
                        A.set(row, col, v); A.set(col, row, v);
                    }
                }
            }
            A.set(k + 1, k, mf_zero<N>()); A.set(k, k + 1, mf_zero<N>());
            k += 2;
        }
    }
    return 0;
}

// ldlt.jl:414-449
template <int N>
MF<N> panel_entry(View<N> A, int64_t i1, int64_t i2, int64_t pfirst, int64_t pivfirst, View<N> dsub, const uint8_t *blocks) {
    int64_t row = std::max(i1, i2), col = std::min(i1, i2);
    MF<N> v = A.get(row, col);
    int64_t q = pfirst;
    while (q < pivfirst) {
        if (blocks[q] == 1) {
            v = sub(v, mul(A.get(row, q), mul(A.get(q, q), A.get(col, q))));
            q += 1;
        } else {
            MF<N> rf = A.get(row, q), rs = A.get(row, q + 1), cf = A.get(col, q), cs = A.get(col, q + 1);
            MF<N> dq = dsub.get(q, 0);
            MF<N> c1 = add(mul(A.get(q, q), cf), mul(dq, cs));
            MF<N> c2 = add(mul(dq, cf), mul(A.get(q + 1, q + 1), cs));
            v = sub(v, add(mul(rf, c1), mul(rs, c2)));
            q += 2;
        }
    }
    return v;
}

// ldlt.jl:451-508
template <int N>
int select_bk_panel(int64_t n, View<N> A, int64_t k, int64_t pfirst, View<N> dsub, const uint8_t *blocks, const MF<N> &alpha, int64_t &pivot) {
    MF<N> diag = panel_entry(A, k, k, pfirst, k, dsub, blocks);
    MF<N> absakk = mf_abs(diag);
    pivot = k;
    if (k == n - 1) return iszero(absakk) ? 0 : 1;
    int64_t imax = k + 1;
    MF<N> colmax = mf_abs(panel_entry(A, imax, k, pfirst, k, dsub, blocks));
    for (int64_t row = k + 2; row < n; ++row) {
        MF<N> cand = mf_abs(panel_entry(A, row, k, pfirst, k, dsub, blocks));
        if (gt(cand, colmax)) { colmax = cand; imax = row; }
    }
    if (eq(mf_max(absakk, colmax), mf_zero<N>())) return 0;
    if (ge(absakk, mul(alpha, colmax))) return 1;
    MF<N> rowmax = mf_zero<N>();
    for (int64_t c = k; c < imax; ++c) rowmax = mf_max(rowmax, mf_abs(panel_entry(A, imax, c, pfirst, k, dsub, blocks)));
    for (int64_t r = imax + 1; r < n; ++r) rowmax = mf_max(rowmax, mf_abs(panel_entry(A, r, imax, pfirst, k, dsub, blocks)));
    if (ge(absakk, mul(mul(alpha, colmax), mf_div(colmax, rowmax)))) return 1;
    MF<N> cd = mf_abs(panel_entry(A, imax, imax, pfirst, k, dsub, blocks));
    pivot = imax;
    return ge(cd, mul(alpha, rowmax)) ? 1 : 2;
}

// ldlt.jl:510-593; returns info (one-based) and panel_last (zero-based)
template <int N>
int64_t ldlt_panel_t(int64_t n, View<N> A, int64_t pfirst, int64_t reqlast, View<N> dsub, int64_t *pivots,
                     uint8_t *blocks, const MF<N> &alpha, int64_t &plast) {
    plast = reqlast;
    int64_t k = pfirst;
    while (k <= plast) {
        int64_t pivot;
        int bs = select_bk_panel(n, A, k, pfirst, dsub, blocks, alpha, pivot);
        if (bs == 0) return k + 1;
        if (bs == 2 && k == plast) {
            if (!(k < n - 1)) return k + 1;
            plast += 1;
        }
        if (bs == 1) {
            pivots[k] = pivot + 1; blocks[k] = 1;
            ldlt_sym_swap(n, A, k, pivot, k);
            MF<N> d = panel_entry(A, k, k, pfirst, k, dsub, blocks);
            if (iszero(d)) return k + 1;
            A.set(k, k, d);
            for (int64_t row = k + 1; row < n; ++row) A.set(row, k, mf_div(panel_entry(A, row, k, pfirst, k, dsub, blocks), d));
            k += 1;
        } else {
            pivots[k] = pivot + 1; blocks[k] = 2; blocks[k + 1] = 0;
            ldlt_sym_swap(n, A, k + 1, pivot, k);
            MF<N> d11 = panel_entry(A, k, k, pfirst, k, dsub, blocks);
            MF<N> d21 = panel_entry(A, k + 1, k, pfirst, k, dsub, blocks);
            MF<N> d22 = panel_entry(A, k + 1, k + 1, pfirst, k, dsub, blocks);
            A.set(k, k, d11); A.set(k + 1, k + 1, d22); dsub.set(k, 0, d21);
            MF<N> s1, s2;
            if (!ldlt_solve_2x2(d11, d21, d22, mf_zero<N>(), mf_zero<N>(), s1, s2)) return k + 1;
            for (int64_t row = k + 2; row < n; ++row) {
                MF<N> f = panel_entry(A, row, k, pfirst, k, dsub, blocks);
                MF<N> s = panel_entry(A, row, k + 1, pfirst, k, dsub, blocks);
                bool ok = ldlt_solve_2x2(d11, d21, d22, f, s, s1, s2);
                if (!(ok && isfin(s1) && isfin(s2))) return k + 1;
                A.set(row, k, s1); A.set(row, k + 1, s2);
            }
            A.set(k + 1, k, mf_zero<N>()); A.set(k, k + 1, mf_zero<N>());
            k += 2;
        }
    }
    return 0;
}

// ldlt_gemmt.jl:1-73 + ldlt.jl:595-632
template <int N>
int64_t ldlt_blocked_t(int64_t n, View<N> A, View<N> dsub, int64_t *pivots, uint8_t *blocks, const MF<N> &alpha,
                       int64_t block, int threads) {
    std::vector<double> wbuf((size_t)N * (size_t)std::max<int64_t>(n, 1) * (size_t)(block + 1));
    const MF<N> one = mf_from<N>(1.0), mone = mf_from<N>(-1.0);
    int64_t pfirst = 0;
    while (pfirst < n) {
        int64_t reqlast = std::min(pfirst + block, n) - 1, plast;
        int64_t info = ldlt_panel_t(n, A, pfirst, reqlast, dsub, pivots, blocks, alpha, plast);
        if (info != 0) return info;
        if (plast < n - 1) {
            const int64_t tfirst = plast + 1, tcount = n - plast - 1, width = plast - pfirst + 1;
            View<N> W{wbuf.data(), tcount, tcount * (block + 1)};
            int64_t q = pfirst;
            while (q <= plast) {
                int64_t lc = q - pfirst;
                if (blocks[q] == 1) {
                    MF<N> d = A.get(q, q);
                    for (int64_t lr = 0; lr < tcount; ++lr) W.set(lr, lc, mul(A.get(tfirst + lr, q), d));
                    q += 1;
                } else {
                    MF<N> d11 = A.get(q, q), d21 = dsub.get(q, 0), d22 = A.get(q + 1, q + 1);
                    for (int64_t lr = 0; lr < tcount; ++lr) {
                        MF<N> f = A.get(tfirst + lr, q), s = A.get(tfirst + lr, q + 1);
                        W.set(lr, lc, add(mul(f, d11), mul(s, d21)));
                        W.set(lr, lc + 1, add(mul(f, d21), mul(s, d22)));
                    }
                    q += 2;
                }
            }
            View<N> L21 = A.sub(tfirst, pfirst), A22 = A.sub(tfirst, tfirst);
            lower_update_t<N>(tcount, width, L21, 0, W, 0, A22, mone, one, threads);
            for (int64_t c = 0; c < tcount; ++c)
                for (int64_t r = c + 1; r < tcount; ++r) A22.set(c, r, A22.get(r, c));
        }
        pfirst = plast + 1;
    }
    return 0;
}

// ldlt.jl:654-717
template <int N>
int64_t sytrf_t(int64_t n, View<N> A, int strategy, int64_t block, int64_t crossover, View<N> dsub, int64_t *pivots,
                uint8_t *blocks, double *origmax, int threads) {
    for (int64_t i = 0; i < n; ++i) { pivots[i] = i + 1; blocks[i] = 0; dsub.set(i, 0, mf_zero<N>()); }
    if (!lower_finite(n, A)) { for (int l = 0; l < N; ++l) origmax[l] = 0.0; return -1; }
    for (int64_t c = 0; c < n; ++c)
        for (int64_t r = c + 1; r < n; ++r) A.set(c, r, A.get(r, c));
    MF<N> mx = mf_zero<N>();
    for (int64_t c = 0; c < n; ++c)
        for (int64_t r = c; r < n; ++r) mx = mf_max(mx, mf_abs(A.get(r, c)));
    for (int l = 0; l < N; ++l) origmax[l] = mx.x[l];
    // ldlt_plan, ldlt.jl:329-350 and ldlt_tuning.jl:82-88
    int64_t b = block > 0 ? block : (N <= 2 ? 16 : (N == 3 ? 12 : 8));
    b = std::max<int64_t>(b, 1);
    bool blocked;
    if (strategy == 1) blocked = false;
    else if (strategy == 2) { blocked = true; b = std::max<int64_t>(b, 2); }
    else blocked = b > 1 && n >= std::max<int64_t>(crossover, 1);
    MF<N> alpha = mf_div(add(mf_from<N>(1.0), mf_sqrt(mf_from<N>(17.0))), mf_from<N>(8.0));
    if (blocked) return ldlt_blocked_t(n, A, dsub, pivots, blocks, alpha, b, threads);
    return ldlt_unblocked_t(n, A, dsub, pivots, blocks, alpha);
}

// ------------------------------------------------------------ solves
// src/solve.jl:23-365. Vector and matrix right-hand sides share arithmetic
// (trsv == one-column trsm bitwise, runtests.jl:683-690).
template <int N> void potrs_t(int64_t n, int64_t nrhs, View<N> F, View<N> B, int threads) {
    const MF<N> one = mf_from<N>(1.0);
    trsm_t<N>(0, 0, 0, 0, n, nrhs, F, B, one, threads);
    trsm_t<N>(0, 0, 1, 0, n, nrhs, F, B, one, threads);
}
template <int N> void getrs_t(int64_t n, int64_t nrhs, View<N> F, const int64_t *ipiv, View<N> B, int threads) {
    const MF<N> one = mf_from<N>(1.0);
    for (int64_t k = 0; k < n; ++k) {
        int64_t p = ipiv[k] - 1;
        if (p != k) for (int64_t c = 0; c < nrhs; ++c) swap_el(B, k, c, p, c);
    }
    trsm_t<N>(0, 0, 0, 1, n, nrhs, F, B, one, threads);
    trsm_t<N>(0, 1, 0, 0, n, nrhs, F, B, one, threads);
}
template <int N> int64_t sytrs_t(int64_t n, int64_t nrhs, View<N> F, View<N> dsub, const int64_t *pivots, const uint8_t *blocks,
                                 View<N> B, int threads) {
    const MF<N> one = mf_from<N>(1.0);
    for (int64_t k = 0; k < n;) {
        if (blocks[k] == 1) { int64_t p = pivots[k] - 1; if (p != k) for (int64_t c = 0; c < nrhs; ++c) swap_el(B, k, c, p, c); k += 1; }
        else if (blocks[k] == 2) { int64_t p = pivots[k] - 1; if (p != k + 1) for (int64_t c = 0; c < nrhs; ++c) swap_el(B, k + 1, c, p, c); k += 2; }
        else return -2;
    }
    trsm_t<N>(0, 0, 0, 1, n, nrhs, F, B, one, threads);
    for (int64_t k = 0; k < n;) {
        if (blocks[k] == 1) {
            MF<N> d = F.get(k, k);
            for (int64_t c = 0; c < nrhs; ++c) B.set(k, c, mf_div(B.get(k, c), d));
            k += 1;
        } else {
            MF<N> d11 = F.get(k, k), d21 = dsub.get(k, 0), d22 = F.get(k + 1, k + 1);
            for (int64_t c = 0; c < nrhs; ++c) {
                MF<N> s1, s2;
                if (!ldlt_solve_2x2(d11, d21, d22, B.get(k, c), B.get(k + 1, c), s1, s2)) return k + 1;
                B.set(k, c, s1); B.set(k + 1, c, s2);
            }
            k += 2;
        }
    }
    trsm_t<N>(0, 0, 1, 1, n, nrhs, F, B, one, threads);
    for (int64_t k = n - 1; k >= 0;) {
        if (blocks[k] == 1) { int64_t p = pivots[k] - 1; if (p != k) for (int64_t c = 0; c < nrhs; ++c) swap_el(B, k, c, p, c); k -= 1; }
        else if (blocks[k] == 0 && k > 0 && blocks[k - 1] == 2) { int64_t s = k - 1; int64_t p = pivots[s] - 1; if (p != s + 1) for (int64_t c = 0; c < nrhs; ++c) swap_el(B, s + 1, c, p, c); k -= 2; }
        else return -2;
    }
    return 0;
}

// -------------------------------------------------- residual / backward error
// src/residual.jl:55-125 (residual!), :127-269 (normwise_backward_error),
// :365-409 (residual_mixed!).
template <int N> MF<N> be_value(MF<N> an, MF<N> xn, MF<N> bn, MF<N> rn, bool finite) {
    if (!finite) return mf_from<N>(std::nan(""));
    MF<N> ps = mf_max(an, bn);
    if (iszero(ps)) return iszero(rn) ? mf_zero<N>() : mf_from<N>(INFINITY);
    MF<N> ss = mf_max(xn, mf_from<N>(1.0));
    MF<N> den = add(mul(mf_div(an, ps), mf_div(xn, ss)), mf_div(mf_div(bn, ps), ss));
    if (iszero(den)) return iszero(rn) ? mf_zero<N>() : mf_from<N>(INFINITY);
    MF<N> sr = mf_div(mf_div(rn, ps), ss);
    return mf_div(sr, den);
}

template <int N> MF<N> mat_el(View<N> A, int uplo, int64_t r, int64_t c) {
    if (uplo == 0) return A.get(r, c);
    if (uplo == 1) return c <= r ? A.get(r, c) : A.get(c, r);
    return c >= r ? A.get(r, c) : A.get(c, r);
}

template <int N>
void backward_error_t(int uplo, int64_t m, int64_t n, int64_t nrhs, View<N> A, View<N> X, View<N> B, View<N> R, double *out) {
    MF<N> an = mf_zero<N>();
    for (int64_t r = 0; r < m; ++r) {
        MF<N> rs = mf_zero<N>();
        for (int64_t c = 0; c < n; ++c) rs = add(rs, mf_abs(mat_el(A, uplo, r, c)));
        an = mf_max(an, rs);
    }
    bool afin = true;
    if (uplo == 0) { for (int64_t c = 0; c < n; ++c) for (int64_t r = 0; r < m; ++r) afin = afin && isfin(A.get(r, c)); }
    else if (uplo == 1) { for (int64_t c = 0; c < n; ++c) for (int64_t r = c; r < m; ++r) afin = afin && isfin(A.get(r, c)); }
    else { for (int64_t c = 0; c < n; ++c) for (int64_t r = 0; r <= c && r < m; ++r) afin = afin && isfin(A.get(r, c)); }
    for (int64_t j = 0; j < nrhs; ++j) {
        bool fin = afin;
        MF<N> xn = mf_zero<N>(), bn = mf_zero<N>(), rn = mf_zero<N>();
        for (int64_t i = 0; i < n; ++i) { fin = fin && isfin(X.get(i, j)); xn = mf_max(xn, mf_abs(X.get(i, j))); }
        for (int64_t i = 0; i < m; ++i) {
            fin = fin && isfin(B.get(i, j)) && isfin(R.get(i, j));
            bn = mf_max(bn, mf_abs(B.get(i, j)));
            rn = mf_max(rn, mf_abs(R.get(i, j)));
        }
        MF<N> v = be_value(an, xn, bn, rn, fin);
        for (int l = 0; l < N; ++l) out[l * nrhs + j] = v.x[l];
    }
}

template <int NS, int NR> MF<NR> promote(const MF<NS> &s) { MF<NR> r = mf_zero<NR>(); for (int l = 0; l < NS; ++l) r.x[l] = s.x[l]; return r; }

template <int NS, int NR>
void residual_mixed_t(int uplo, int64_t m, int64_t n, int64_t nrhs, View<NS> A, View<NS> X, View<NS> B, View<NR> R, int threads) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads_or(threads))
    for (int64_t j = 0; j < nrhs; ++j)
        for (int64_t r = 0; r < m; ++r) {
            MF<NR> acc = promote<NS, NR>(B.get(r, j));
            for (int64_t c = 0; c < n; ++c) acc = msub(acc, promote<NS, NR>(mat_el(A, uplo, r, c)), promote<NS, NR>(X.get(c, j)));
            R.set(r, j, acc);
        }
}

template <int N> MF<N> fused_or_fail(const MF<N> &c, const MF<N> &a, const MF<N> &b, bool &ok) {
    if constexpr (N == 3) { ok = true; return mulacc_x3(c, a, b); }
    ok = false; return c;
}

template <int N> int arith_t(int op, int64_t count, const double *x, const double *y, const double *z, double *out) {
    for (int64_t i = 0; i < count; ++i) {
        MF<N> a, b, c, r;
        for (int l = 0; l < N; ++l) { a.x[l] = x[l * count + i]; b.x[l] = y ? y[l * count + i] : 0.0; c.x[l] = z ? z[l * count + i] : 0.0; }
        bool ok = true;
        switch (op) {
            case 0: r = add(a, b); break;
            case 1: r = sub(a, b); break;
            case 2: r = mul(a, b); break;
            case 3: r = mf_div(a, b); break;
            case 4: r = mf_sqrt(a); break;
            case 5: r = mac(c, a, b); break;
            case 6: r = fused_or_fail(c, a, b, ok); if (!ok) return -101; break;
            case 7: r = mf_inv(a); break;
            case 8: r = msub(c, a, b); break;
            default: return -102;
        }
        for (int l = 0; l < N; ++l) out[l * count + i] = r.x[l];
    }
    return 0;
}

}  // namespace

// ======================================================================
// extern "C" surface for oracle/oracle.py (ctypes)
// ======================================================================
// ------------------------------------------------------------------- QR
// src/factorizations/qr.jl. Column norms are kept as a scale-safe pair
// (scale, scaled_sum) with norm = scale * sqrt(scaled_sum) (qr.jl:87-165).
// eps(Float64xN) = eps(Float64)^N = 2^-52N (MultiFloats' definition), so the
// pivot margin sqrt(eps) = 2^-26N and the downdate floor 16 eps = 2^(4-52N)
// (qr.jl:721-722) are exact powers of two.
template <int N> struct NormState { MF<N> scale, sum; };

template <int N> static inline bool le(const MF<N> &x, const MF<N> &y) { return lt(x, y) || eq(x, y); }

// qr.jl:96-122 (value_at(row) = A[row, col]) and qr.jl:229-262 (the delayed
// variant: value_at(row) = A[row, col] - sum_prior A[row, bs+prior] F[prior, col]).
template <int N, class Value>
NormState<N> qr_norm_scan(int64_t first_row, int64_t m, Value value_at) {
    const MF<N> one = mf_from<N>(1.0);
    MF<N> scale = mf_zero<N>(), ssum = one;
    bool seen = false;
    for (int64_t row = first_row; row < m; ++row) {
        const MF<N> v = mf_abs(value_at(row));
        if (iszero(v)) continue;
        if (!seen) { scale = v; ssum = one; seen = true; }
        else if (gt(v, scale)) { const MF<N> r = mf_div(scale, v); ssum = add(one, mul(mul(ssum, r), r)); scale = v; }
        else { const MF<N> r = mf_div(v, scale); ssum = add(ssum, mul(r, r)); }
    }
    return {scale, ssum};
}

template <int N> NormState<N> qr_col_state(int64_t m, View<N> A, int64_t first_row, int64_t col) {
    return qr_norm_scan<N>(first_row, m, [&](int64_t r) { return A.get(r, col); });
}

template <int N> NormState<N> qr_delayed_state(int64_t m, View<N> A, View<N> F, int64_t bs, int64_t pc, int64_t first_row, int64_t col) {
    return qr_norm_scan<N>(first_row, m, [&](int64_t r) {
        MF<N> v = A.get(r, col);
        for (int64_t p = 0; p < pc; ++p) v = sub(v, mul(A.get(r, bs + p), F.get(p, col)));
        return v;
    });
}

template <int N> MF<N> qr_norm_of(const NormState<N> &st) { return iszero(st.scale) ? mf_zero<N>() : mul(st.scale, mf_sqrt(st.sum)); }

// qr.jl:606-622: Householder reflector of column step; returns tau.
template <int N> MF<N> qr_make_reflector(int64_t m, View<N> A, int64_t step) {
    const MF<N> nrm = qr_norm_of(qr_col_state(m, A, step, step));
    if (iszero(nrm)) return mf_zero<N>();
    const MF<N> alpha = A.get(step, step);
    const MF<N> beta = ge(alpha, mf_zero<N>()) ? mf_neg(nrm) : nrm;
    const MF<N> leading = sub(alpha, beta);
    const MF<N> tau = mf_div(sub(beta, alpha), beta);
    for (int64_t row = step + 1; row < m; ++row) A.set(row, step, mf_div(A.get(row, step), leading));
    A.set(step, step, beta);
    return tau;
}

// qr.jl:624-642 (and the apply_q! form qr.jl:890-929): column col of T gets
// H = I - tau v v' with v = (1, V[step+1:m, vcol]).
template <int N> void qr_reflect_column(int64_t m, View<N> V, int64_t vcol, int64_t step, const MF<N> &tau, View<N> T, int64_t col) {
    MF<N> proj = T.get(step, col);
    for (int64_t row = step + 1; row < m; ++row) proj = add(proj, mul(V.get(row, vcol), T.get(row, col)));
    proj = mul(proj, tau);
    T.set(step, col, sub(T.get(step, col), proj));
    for (int64_t row = step + 1; row < m; ++row) T.set(row, col, sub(T.get(row, col), mul(V.get(row, vcol), proj)));
}

// qr.jl:644-669
template <int N> void qr_apply_trailing(int64_t m, int64_t n, View<N> A, const MF<N> &tau, int64_t step) {
    if (iszero(tau)) return;
    for (int64_t c = step + 1; c < n; ++c) qr_reflect_column(m, A, step, step, tau, A, c);
}

template <int N> struct QrState {
    std::vector<MF<N>> scale, sum;
    std::vector<char> dirty;
};

// qr.jl:167-182
template <int N> MF<N> qr_recompute(int64_t m, View<N> A, QrState<N> &S, int64_t first_row, int64_t col) {
    const NormState<N> st = qr_col_state(m, A, first_row, col);
    S.scale[col] = st.scale; S.sum[col] = st.sum; S.dirty[col] = 0;
    return qr_norm_of(st);
}

template <int N> void qr_swap(int64_t m, View<N> A, int64_t *perm, QrState<N> &S, int64_t a, int64_t b) {
    if (a == b) return;
    for (int64_t r = 0; r < m; ++r) { const MF<N> t = A.get(r, a); A.set(r, a, A.get(r, b)); A.set(r, b, t); }
    std::swap(perm[a], perm[b]);
    std::swap(S.scale[a], S.scale[b]); std::swap(S.sum[a], S.sum[b]); std::swap(S.dirty[a], S.dirty[b]);
}

// qr.jl:264-301 (F empty: pc = 0)
template <int N> void qr_update_norms(int64_t m, int64_t n, View<N> A, const MF<N> &tau, int64_t step, QrState<N> &S, const MF<N> &floor_,
                                      View<N> F, int64_t bs, int64_t pc) {
    if (iszero(tau)) return;
    for (int64_t c = step + 1; c < n; ++c) {
        if (S.dirty[c]) continue;
        const MF<N> cs = S.scale[c], removed = mf_abs(A.get(step, c));
        if (iszero(cs)) { S.dirty[c] = !iszero(removed); continue; }
        const MF<N> ratio = mf_div(removed, cs);
        const MF<N> ns = sub(S.sum[c], mul(ratio, ratio));
        if (!isfin(ns) || le(ns, floor_)) {
            const NormState<N> st = pc > 0 ? qr_delayed_state(m, A, F, bs, pc, step + 1, c) : qr_col_state(m, A, step + 1, c);
            S.scale[c] = st.scale; S.sum[c] = st.sum; S.dirty[c] = 0;
        } else {
            S.sum[c] = ns;
        }
    }
}

// qr.jl:763-799: unpivoted Householder QR; tau has min(m, n) entries.
template <int N> void geqrf_t(int64_t m, int64_t n, View<N> A, View<N> tau) {
    const int64_t k = std::min(m, n);
    for (int64_t step = 0; step < k; ++step) {
        const MF<N> t = qr_make_reflector(m, A, step);
        tau.set(step, 0, t);
        qr_apply_trailing(m, n, A, t, step);
    }
}

// qr.jl:192-224
template <int N> int64_t qr_pivot_hybrid(int64_t m, int64_t n, View<N> A, const int64_t *perm, int64_t step, QrState<N> &S, const MF<N> &onem) {
    int64_t pivot = step;
    MF<N> pn = qr_recompute(m, A, S, step, step);
    for (int64_t c = step + 1; c < n; ++c) {
        MF<N> cand = qr_norm_of(NormState<N>{S.scale[c], S.sum[c]});
        if (S.dirty[c] || iszero(pn) || ge(cand, mul(onem, pn))) cand = qr_recompute(m, A, S, step, c);
        if (gt(cand, pn) || (eq(cand, pn) && perm[c] < perm[pivot])) { pivot = c; pn = cand; }
    }
    return pivot;
}

// qr.jl:344-362
template <int N> int64_t qr_pivot_downdated(int64_t n, const int64_t *perm, int64_t step, const QrState<N> &S) {
    int64_t pivot = step;
    MF<N> pn = qr_norm_of(NormState<N>{S.scale[step], S.sum[step]});
    for (int64_t c = step + 1; c < n; ++c) {
        const MF<N> cand = qr_norm_of(NormState<N>{S.scale[c], S.sum[c]});
        if (gt(cand, pn) || (eq(cand, pn) && perm[c] < perm[pivot])) { pivot = c; pn = cand; }
    }
    return pivot;
}

// qr.jl:692-743 with the unblocked route (qr.jl:364-398) and the blocked
// DLAQPS route (qr.jl:450-604, blocked when min(m, n) >= 16 and m*n >= 16384,
// qr.jl:303-314). perm is one-based. Returns 0 or -1 (non-finite input: A,
// tau and perm untouched apart from tau = 0, perm = identity).
template <int N> int64_t geqp3_t(int64_t m, int64_t n, View<N> A, View<N> tau, int64_t *perm, int route) {
    const int64_t k = std::min(m, n);
    for (int64_t c = 0; c < n; ++c) perm[c] = c + 1;
    for (int64_t s = 0; s < k; ++s) tau.set(s, 0, mf_zero<N>());
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i)
            if (!isfin(A.get(i, j))) return -1;
    QrState<N> S;
    S.scale.resize(n); S.sum.resize(n); S.dirty.assign(n, 0);
    for (int64_t c = 0; c < n; ++c) { const NormState<N> st = qr_col_state(m, A, 0, c); S.scale[c] = st.scale; S.sum[c] = st.sum; }
    const MF<N> one = mf_from<N>(1.0), mone = mf_from<N>(-1.0);
    const MF<N> onem = sub(one, mf_from<N>(std::ldexp(1.0, -26 * N)));
    const MF<N> floor_ = mf_from<N>(std::ldexp(1.0, 4 - 52 * N));
    const bool blocked = route == 2 || (route == 0 && k >= 16 && m * n >= 16384);
    if (!blocked) {
        View<N> none{nullptr, 0, 0};
        for (int64_t step = 0; step < k; ++step) {
            const int64_t pivot = qr_pivot_hybrid(m, n, A, perm, step, S, onem);
            qr_swap(m, A, perm, S, step, pivot);
            const MF<N> t = qr_make_reflector(m, A, step);
            tau.set(step, 0, t);
            qr_apply_trailing(m, n, A, t, step);
            qr_update_norms(m, n, A, t, step, S, floor_, none, step, 0);
        }
        return 0;
    }
    const int64_t bsz = std::min<int64_t>(16, k);
    std::vector<double> Fbuf((size_t)N * bsz * n), auxbuf((size_t)N * bsz), ybuf((size_t)N * n);
    View<N> F{Fbuf.data(), bsz, bsz * n}, aux{auxbuf.data(), bsz, bsz};
    int64_t bs = 0;
    while (bs < k) {
        const int64_t requested = std::min(bsz, k - bs);
        std::fill(Fbuf.begin(), Fbuf.end(), 0.0);
        int64_t actual = 0;
        for (int64_t ls = 0; ls < requested; ++ls) {
            const int64_t step = bs + ls;
            const int64_t pivot = qr_pivot_downdated(n, perm, step, S);
            if (pivot != step) {
                for (int64_t r = 0; r < m; ++r) { const MF<N> t = A.get(r, step); A.set(r, step, A.get(r, pivot)); A.set(r, pivot, t); }
                for (int64_t p = 0; p < ls; ++p) { const MF<N> t = F.get(p, step); F.set(p, step, F.get(p, pivot)); F.set(p, pivot, t); }
                std::swap(perm[step], perm[pivot]);
                std::swap(S.scale[step], S.scale[pivot]); std::swap(S.sum[step], S.sum[pivot]); std::swap(S.dirty[step], S.dirty[pivot]);
            }
            if (ls > 0)
                for (int64_t r = step; r < m; ++r) {
                    MF<N> corr = mf_zero<N>();
                    for (int64_t p = 0; p < ls; ++p) corr = add(corr, mul(A.get(r, bs + p), F.get(p, step)));
                    A.set(r, step, sub(A.get(r, step), corr));
                }
            const MF<N> t = qr_make_reflector(m, A, step);
            tau.set(step, 0, t);
            const MF<N> diagonal = A.get(step, step);
            A.set(step, step, one);
            if (step < n - 1) {   // F[ls, step+1:n] = tau * A[step:m, step+1:n]' v   (gemv!, trans = :T)
                const int64_t cols = n - step - 1;
                View<N> y{ybuf.data(), 1, cols};
                gemv_t<N>(1, m - step, cols, A.sub(step, step + 1), A.sub(step, step), y, t, mf_zero<N>());
                for (int64_t o = 0; o < cols; ++o) F.set(ls, step + 1 + o, y.get(o, 0));
            }
            if (ls > 0) {
                gemv_t<N>(1, m - step, ls, A.sub(step, bs), A.sub(step, step), View<N>{aux.p, 1, bsz}, mf_neg(t), mf_zero<N>());
                for (int64_t c = step + 1; c < n; ++c) {
                    MF<N> corr = mf_zero<N>();
                    for (int64_t p = 0; p < ls; ++p) corr = add(corr, mul(F.get(p, c), aux.get(p, 0)));
                    F.set(ls, c, add(F.get(ls, c), corr));
                }
            }
            for (int64_t c = step + 1; c < n; ++c) {
                MF<N> corr = mf_zero<N>();
                for (int64_t p = 0; p <= ls; ++p) corr = add(corr, mul(A.get(step, bs + p), F.get(p, c)));
                A.set(step, c, sub(A.get(step, c), corr));
            }
            A.set(step, step, diagonal);
            qr_update_norms(m, n, A, t, step, S, floor_, F, bs, ls + 1);
            actual = ls + 1;
            bool unreliable = false;
            for (int64_t c = step + 1; c < n; ++c) unreliable = unreliable || S.dirty[c];
            if (unreliable) break;
        }
        const int64_t be = bs + actual - 1;
        if (be + 1 < m && be + 1 < n)
            gemm_t<N>(m - be - 1, n - be - 1, actual, A.sub(be + 1, bs), View<N>{F.p + (be + 1) * bsz, bsz, bsz * n}, A.sub(be + 1, be + 1), mone, one,
                      true, 1);
        for (int64_t c = be + 1; c < n; ++c) qr_recompute(m, A, S, be + 1, c);
        bs = be + 1;
    }
    return 0;
}

// qr.jl:938-956: B <- Q B (trans = 0, reflectors in reverse) or Q' B.
template <int N> void ormqr_t(int trans, int64_t m, int64_t k, View<N> Fq, View<N> tau, int64_t nrhs, View<N> B) {
    for (int64_t i = 0; i < k; ++i) {
        const int64_t step = trans ? i : k - 1 - i;
        const MF<N> t = tau.get(step, 0);
        if (iszero(t)) continue;
        for (int64_t c = 0; c < nrhs; ++c) qr_reflect_column(m, Fq, step, step, t, B, c);
    }
}

#define DISPATCH(N, ...) \
    switch (N) { case 1: { constexpr int NN = 1; __VA_ARGS__; } break; case 2: { constexpr int NN = 2; __VA_ARGS__; } break; \
                 case 3: { constexpr int NN = 3; __VA_ARGS__; } break; case 4: { constexpr int NN = 4; __VA_ARGS__; } break; default: return -100; }

extern "C" {

// Elementwise arithmetic over `count` elements stored planar (limb stride = count).
// op: 0 add, 1 sub, 2 mul, 3 div, 4 sqrt(x), 5 mac z + x*y, 6 fused x3 mulacc z + x*y,
//     7 inv(x), 8 msub z - x*y
int orc_arith(int N, int op, int64_t count, const double *x, const double *y, const double *z, double *out) {
    DISPATCH(N, return arith_t<NN>(op, count, x, y, z, out));
    return 0;
}

int orc_two_sum(int64_t count, const double *a, const double *b, double *s, double *e, int fast) {
    for (int64_t i = 0; i < count; ++i) { if (fast) fast_two_sum(a[i], b[i], s[i], e[i]); else two_sum(a[i], b[i], s[i], e[i]); }
    return 0;
}
int orc_two_prod(int64_t count, const double *a, const double *b, double *p, double *e) {
    for (int64_t i = 0; i < count; ++i) two_prod(a[i], b[i], p[i], e[i]);
    return 0;
}

int orc_gemm(int N, int64_t m, int64_t n, int64_t k, double *A, int64_t lda, int64_t lsa, double *B, int64_t ldb, int64_t lsb,
             double *C, int64_t ldc, int64_t lsc, const double *alpha, const double *beta, int fused, int threads) {
    DISPATCH(N, gemm_t<NN>(m, n, k, View<NN>{A, lda, lsa}, View<NN>{B, ldb, lsb}, View<NN>{C, ldc, lsc},
                           scal<NN>(alpha), scal<NN>(beta), fused != 0, threads));
    return 0;
}

int orc_lower_update(int N, int64_t n, int64_t k, double *L, int64_t ldl, int64_t lsl, int ltrans, double *R, int64_t ldr,
                     int64_t lsr, int rtrans, double *C, int64_t ldc, int64_t lsc, const double *alpha, const double *beta, int threads) {
    DISPATCH(N, lower_update_t<NN>(n, k, View<NN>{L, ldl, lsl}, ltrans, View<NN>{R, ldr, lsr}, rtrans, View<NN>{C, ldc, lsc},
                                   scal<NN>(alpha), scal<NN>(beta), threads));
    return 0;
}

int orc_trsm(int N, int side_right, int upper, int trans, int unit, int64_t m, int64_t n, double *A, int64_t lda, int64_t lsa,
             double *B, int64_t ldb, int64_t lsb, const double *alpha, int threads) {
    DISPATCH(N, trsm_t<NN>(side_right, upper, trans, unit, m, n, View<NN>{A, lda, lsa}, View<NN>{B, ldb, lsb}, scal<NN>(alpha), threads));
    return 0;
}

int orc_trmm(int N, int side_right, int upper, int trans, int unit, int64_t m, int64_t n, double *A, int64_t lda, int64_t lsa,
             double *B, int64_t ldb, int64_t lsb, const double *alpha) {
    DISPATCH(N, trmm_t<NN>(side_right, upper, trans, unit, m, n, View<NN>{A, lda, lsa}, View<NN>{B, ldb, lsb}, scal<NN>(alpha)));
    return 0;
}

int orc_dot(int N, int64_t m, int64_t n, double *x, int64_t ldx, int64_t lsx, double *y, int64_t ldy, int64_t lsy, double *out) {
    DISPATCH(N, { const MF<NN> r = dot_t<NN>(m, n, View<NN>{x, ldx, lsx}, View<NN>{y, ldy, lsy}); for (int l = 0; l < NN; ++l) out[l] = r.x[l]; });
    return 0;
}

int orc_syrk_packed(int N, int64_t n, int64_t rf, int64_t rl, double *P, int64_t ldp, int64_t lsp, double *out, int64_t ls,
                    const double *alpha, const double *beta) {
    DISPATCH(N, syrk_packed_t<NN>(n, rf, rl, View<NN>{P, ldp, lsp}, out, ls, scal<NN>(alpha), scal<NN>(beta)));
    return 0;
}

int orc_gemv(int N, int trans, int64_t m, int64_t n, double *A, int64_t lda, int64_t lsa, double *x, int64_t lsx, double *y,
             int64_t lsy, const double *alpha, const double *beta) {
    DISPATCH(N, gemv_t<NN>(trans, m, n, View<NN>{A, lda, lsa}, View<NN>{x, 1, lsx}, View<NN>{y, 1, lsy}, scal<NN>(alpha), scal<NN>(beta)));
    return 0;
}

int orc_symv(int N, int upper, int64_t n, double *A, int64_t lda, int64_t lsa, double *x, int64_t lsx, double *y, int64_t lsy,
             const double *alpha, const double *beta) {
    DISPATCH(N, symv_t<NN>(upper, n, View<NN>{A, lda, lsa}, View<NN>{x, 1, lsx}, View<NN>{y, 1, lsy}, scal<NN>(alpha), scal<NN>(beta)));
    return 0;
}

int64_t orc_potrf(int N, int64_t n, double *A, int64_t lda, int64_t lsa, int64_t block, int threads) {
    DISPATCH(N, return potrf_t<NN>(n, View<NN>{A, lda, lsa}, block, threads));
    return 0;
}

int64_t orc_getrf(int N, int64_t m, int64_t n, double *A, int64_t lda, int64_t lsa, int64_t block, int64_t *ipiv, double *origmax,
                  int gemm_fused, int threads) {
    DISPATCH(N, return getrf_t<NN>(m, n, View<NN>{A, lda, lsa}, block, ipiv, origmax, gemm_fused, threads));
    return 0;
}

int64_t orc_sytrf(int N, int64_t n, double *A, int64_t lda, int64_t lsa, int strategy, int64_t block, int64_t crossover,
                  double *dsub, int64_t lsd, int64_t *pivots, uint8_t *blocks, double *origmax, int threads) {
    DISPATCH(N, return sytrf_t<NN>(n, View<NN>{A, lda, lsa}, strategy, block, crossover, View<NN>{dsub, 1, lsd}, pivots, blocks, origmax, threads));
    return 0;
}

int orc_potrs(int N, int64_t n, int64_t nrhs, double *F, int64_t ldf, int64_t lsf, double *B, int64_t ldb, int64_t lsb, int threads) {
    DISPATCH(N, potrs_t<NN>(n, nrhs, View<NN>{F, ldf, lsf}, View<NN>{B, ldb, lsb}, threads));
    return 0;
}

int orc_getrs(int N, int64_t n, int64_t nrhs, double *F, int64_t ldf, int64_t lsf, const int64_t *ipiv, double *B, int64_t ldb,
              int64_t lsb, int threads) {
    DISPATCH(N, getrs_t<NN>(n, nrhs, View<NN>{F, ldf, lsf}, ipiv, View<NN>{B, ldb, lsb}, threads));
    return 0;
}

int64_t orc_sytrs(int N, int64_t n, int64_t nrhs, double *F, int64_t ldf, int64_t lsf, double *dsub, int64_t lsd, const int64_t *pivots,
                  const uint8_t *blocks, double *B, int64_t ldb, int64_t lsb, int threads) {
    DISPATCH(N, return sytrs_t<NN>(n, nrhs, View<NN>{F, ldf, lsf}, View<NN>{dsub, 1, lsd}, pivots, blocks, View<NN>{B, ldb, lsb}, threads));
    return 0;
}

int orc_backward_error(int N, int uplo, int64_t m, int64_t n, int64_t nrhs, double *A, int64_t lda, int64_t lsa, double *X, int64_t ldx,
                       int64_t lsx, double *B, int64_t ldb, int64_t lsb, double *R, int64_t ldr, int64_t lsr, double *out) {
    DISPATCH(N, backward_error_t<NN>(uplo, m, n, nrhs, View<NN>{A, lda, lsa}, View<NN>{X, ldx, lsx}, View<NN>{B, ldb, lsb},
                                     View<NN>{R, ldr, lsr}, out));
    return 0;
}

int orc_residual_mixed(int NS, int NR, int uplo, int64_t m, int64_t n, int64_t nrhs, double *A, int64_t lda, int64_t lsa, double *X,
                       int64_t ldx, int64_t lsx, double *B, int64_t ldb, int64_t lsb, double *R, int64_t ldr, int64_t lsr, int threads) {
#define RM(S, T) residual_mixed_t<S, T>(uplo, m, n, nrhs, View<S>{A, lda, lsa}, View<S>{X, ldx, lsx}, View<S>{B, ldb, lsb}, View<T>{R, ldr, lsr}, threads)
    if (NS == 2 && NR == 3) RM(2, 3);
    else if (NS == 2 && NR == 4) RM(2, 4);
    else if (NS == 3 && NR == 4) RM(3, 4);
    else if (NS == 1 && NR == 4) RM(1, 4);
    else if (NS == 1 && NR == 2) RM(1, 2);
    else if (NS == 1 && NR == 3) RM(1, 3);
    else return -100;
#undef RM
    return 0;
}

int orc_threads(void) { return nthreads_or(0); }

int orc_geqrf(int N, int64_t m, int64_t n, double *A, int64_t lda, int64_t lsa, double *tau, int64_t lst) {
    DISPATCH(N, geqrf_t<NN>(m, n, View<NN>{A, lda, lsa}, View<NN>{tau, 1, lst}));
    return 0;
}

// route: 0 MFLA's choice, 1 unblocked, 2 blocked
int64_t orc_geqp3(int N, int64_t m, int64_t n, double *A, int64_t lda, int64_t lsa, double *tau, int64_t lst, int64_t *perm, int route) {
    DISPATCH(N, return geqp3_t<NN>(m, n, View<NN>{A, lda, lsa}, View<NN>{tau, 1, lst}, perm, route));
    return 0;
}

int orc_ormqr(int N, int trans, int64_t m, int64_t k, double *F, int64_t ldf, int64_t lsf, double *tau, int64_t lst, int64_t nrhs, double *B,
              int64_t ldb, int64_t lsb) {
    DISPATCH(N, ormqr_t<NN>(trans, m, k, View<NN>{F, ldf, lsf}, View<NN>{tau, 1, lst}, nrhs, View<NN>{B, ldb, lsb}));
    return 0;
}

}  // extern "C"
