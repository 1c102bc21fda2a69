// Factorization kernels for Float64xN: Cholesky panel, partial-pivoting LU
// panel steps, Bunch-Kaufman LDLT (unblocked and lazy blocked panel), and the
// pivot / block-diagonal steps of the solves.
//
// All kernels share the call's device Status word: they return immediately
// once a previous kernel recorded a failure, so the matrix is left exactly as
// MFLA leaves it when the factorization stops.
#include <type_traits>
#include "mf_common.hpp"
#include "k_factor.hpp"
#include "k_factor_dev.hpp"

namespace mfk {

// -------------------------------------------------------------- Cholesky
// src/factorizations/cholesky.jl:10-34, one workgroup per panel diagonal
// block. Per column c: thread 0 forms d = A[c,c] - sum_k A[c,k]^2 while
// thread t forms v = A[c+t,c] - sum_k A[c+t,k] A[c,k] for the rows below,
// both over k ascending from the panel start (the same element formulas and
// order as MFLA, evaluated concurrently); then sqrt(d) and v / sqrt(d).
// Panels of up to 32 columns keep their lower triangle in LDS.
constexpr int CHOL_LDS = 32;

// Diagonal block of one panel (cholesky.jl:10-34), nb <= 32, in LDS. Every
// element (r, c), r >= c, keeps its running value v = A(r, c) - sum_k
// L(r, k) L(c, k) in a register and receives term k as soon as column k is
// final, so a column's critical path is one msub, the square root and the
// division instead of its whole chain; each element still applies its terms
// in ascending k, then the square root (diagonal) or the division by the
// pivot (below it), exactly MFLA's sequence.
// PB: the largest panel width of the instance (16: the default 16-column
// panels, whose 8 KB of LDS fit beside the look-ahead update's workgroups).
template <int PB> constexpr int chol_ept() { return (PB * (PB + 1) / 2 + 255) / 256; }   // elements per thread

template <int N, int PB>
__global__ __launch_bounds__(256, 8) void k_chol_panel_lds(DV A, int64_t first, int64_t last, Status *st) {
    __builtin_amdgcn_s_setprio(3);   // runs beside the look-ahead SYRK: 64 VGPRs, issued first
    if (st->info != 0) return;
    constexpr int CHOL_EPT = chol_ept<PB>();
    __shared__ double P[PB * PB * N];     // P[(q * PB + r) * N + l] = L(first + r, first + q)
    __shared__ double piv_s[N];
    __shared__ int fail_s;
    const int nb = (int)(last - first + 1);
    const int t = threadIdx.x;
    const int count = nb * (nb + 1) / 2;
    int er[CHOL_EPT], ec[CHOL_EPT];
    mf::V<N> v[CHOL_EPT];
#pragma unroll
    for (int i = 0; i < CHOL_EPT; ++i) {
        const int e = t + 256 * i;
        er[i] = -1; ec[i] = -1;
        if (e < count) {                               // packed lower triangle, by columns
            int c = 0, start = 0;
            while (start + (nb - c) <= e) { start += nb - c; ++c; }
            ec[i] = c;
            er[i] = c + (e - start);
            v[i] = ld<N>(A, first + er[i], first + c);
        } else {
            v[i] = mf::zero<N>();
        }
    }
    auto get = [&](int r, int q) {
        mf::V<N> x;
#pragma unroll
        for (int l = 0; l < N; ++l) x.x[l] = P[(q * PB + r) * N + l];
        return x;
    };
    for (int k = 0; k < nb; ++k) {
#pragma unroll
        for (int i = 0; i < CHOL_EPT; ++i) {
            if (er[i] == k && ec[i] == k) {            // the pivot
                if (!mf::gt(v[i], mf::zero<N>())) {
                    fail_s = 1;
                    st->info = first + k + 1;
                } else {
                    fail_s = 0;
                    const mf::V<N> p = mf::sqrt(v[i]);
                    stv<N>(A, first + k, first + k, p);
#pragma unroll
                    for (int l = 0; l < N; ++l) { piv_s[l] = p.x[l]; P[(k * PB + k) * N + l] = p.x[l]; }
                }
            }
        }
        __syncthreads();
        if (fail_s) return;
        mf::V<N> p;
#pragma unroll
        for (int l = 0; l < N; ++l) p.x[l] = piv_s[l];
#pragma unroll
        for (int i = 0; i < CHOL_EPT; ++i) {
            if (ec[i] == k && er[i] > k) {             // column k below the pivot
                const mf::V<N> q = mf::div(v[i], p);
                stv<N>(A, first + er[i], first + k, q);
#pragma unroll
                for (int l = 0; l < N; ++l) P[(k * PB + er[i]) * N + l] = q.x[l];
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < CHOL_EPT; ++i)             // term k of every later element
            if (ec[i] > k) v[i] = mf::msub(v[i], get(er[i], k), get(ec[i], k));
    }
}

// Panels wider than 32 columns: the same computation from global memory.
template <int N>
__global__ __launch_bounds__(256) void k_chol_panel(DV A, int64_t first, int64_t last, Status *st) {
    if (st->info != 0) return;
    __shared__ double piv_s[N];
    __shared__ int fail_s;
    for (int64_t c = first; c <= last; ++c) {
        const int64_t r = c + threadIdx.x;
        mf::V<N> v = mf::zero<N>();
        if (r <= last) {
            v = ld<N>(A, r, c);
            for (int64_t k = first; k < c; ++k) v = mf::msub(v, ld<N>(A, r, k), ld<N>(A, c, k));
        }
        if (threadIdx.x == 0) {
            if (!mf::gt(v, mf::zero<N>())) {
                fail_s = 1;
                st->info = c + 1;
            } else {
                fail_s = 0;
                mf::V<N> p = mf::sqrt(v);
                stv<N>(A, c, c, p);
                lds_put<N>(piv_s, p);
            }
        }
        __syncthreads();
        if (fail_s) return;
        const mf::V<N> p = lds_get<N>(piv_s);
        for (int64_t rr = c + threadIdx.x; rr <= last; rr += blockDim.x) {
            if (rr == c) continue;
            mf::V<N> w = (rr == r) ? v : ld<N>(A, rr, c);
            if (rr != r)
                for (int64_t k = first; k < c; ++k) w = mf::msub(w, ld<N>(A, rr, k), ld<N>(A, c, k));
            stv<N>(A, rr, c, mf::div(w, p));
        }
        __syncthreads();
    }
}

// -------------------------------------------------------------------- LU
// lu.jl:91-102: pivot search over rows k..m-1 of column k, record ipiv, then
// swap the rows within the panel columns [c0, c1) (MFLA swaps whole rows;
// the other columns get the same interchanges from k_lu_laswp once the panel
// is done, before anything reads them, so every element ends identical).
// One 256-thread workgroup. Each thread scans its rows in ascending order
// (chunks of 8 rows whose loads are issued together) keeping the first
// strict maximum of |a|; the waves then reduce lexicographically -- for
// each limb the maximum over the lanes still tied on the higher limbs (DPP
// row rotations and lane reads), then the smallest row among the lanes left
// -- which is MFLA's first strict maximum for finite values. A NaN in row k
// keeps the pivot at k, as MFLA's scan does.
template <int CTRL>
__device__ __forceinline__ double lp_dpp(double x) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lp_rl(double x, int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), l), __builtin_amdgcn_readlane(__double2loint(x), l));
}
__device__ __forceinline__ double lp_wave_max(double x) {
    x = fmax(x, lp_dpp<0xB1>(x));
    x = fmax(x, lp_dpp<0x4E>(x));
    x = fmax(x, lp_dpp<0x124>(x));
    x = fmax(x, lp_dpp<0x128>(x));
    return fmax(fmax(lp_rl(x, 0), lp_rl(x, 16)), fmax(lp_rl(x, 32), lp_rl(x, 48)));
}
__device__ __forceinline__ int lp_wave_min(int i) {
    i = min(i, __builtin_amdgcn_update_dpp(0, i, 0xB1, 0xf, 0xf, false));
    i = min(i, __builtin_amdgcn_update_dpp(0, i, 0x4E, 0xf, 0xf, false));
    i = min(i, __builtin_amdgcn_update_dpp(0, i, 0x124, 0xf, 0xf, false));
    i = min(i, __builtin_amdgcn_update_dpp(0, i, 0x128, 0xf, 0xf, false));
    return min(min(__builtin_amdgcn_readlane(i, 0), __builtin_amdgcn_readlane(i, 16)),
               min(__builtin_amdgcn_readlane(i, 32), __builtin_amdgcn_readlane(i, 48)));
}

// lexicographic wave argmax: every lane ends with the wave's best value and
// its smallest row
template <int N>
__device__ __forceinline__ void lp_wave_argmax(mf::V<N> &best, int &bi) {
    bool cand = true;
    mf::V<N> m;
#pragma unroll
    for (int l = 0; l < N; ++l) {
        m.x[l] = lp_wave_max(cand ? best.x[l] : -INFINITY);
        cand = cand && best.x[l] == m.x[l];
    }
    bi = lp_wave_min(cand ? bi : INT32_MAX);
    best = m;
}

// The panel kernels run while the look-ahead trailing update fills the CUs:
// they are held to 64 VGPRs (8 waves per SIMD), which fit beside the update's
// waves, and raise their wave priority so the SIMD issues their
// instructions first.
constexpr int LP_THREADS = 256, LP_CHUNK = 4;

template <int N>
__global__ __launch_bounds__(LP_THREADS, 8) void k_lu_pivot_swap(DV A, int64_t k, int64_t m, int64_t c0, int64_t c1, int64_t *ipiv, Status *st) {
    __builtin_amdgcn_s_setprio(3);
    __shared__ double wv[LP_THREADS / 64][N];
    __shared__ int wi[LP_THREADS / 64];
    __shared__ int nan_kk;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int64_t info = st->info;                 // checked after the first loads are in flight
    const int rows = (int)(m - k);
    mf::V<N> best = mf::zero<N>();
    best.x[0] = -1.0;                              // below every |a|; a NaN never replaces it
    int bi = INT32_MAX;
    for (int base = 0; base < rows; base += LP_THREADS * LP_CHUNK) {
        mf::V<N> v[LP_CHUNK];
#pragma unroll
        for (int i = 0; i < LP_CHUNK; ++i) {       // clamped rows: every load in bounds
            const int r = base + t + LP_THREADS * i;
            v[i] = ld<N>(A, k + (r < rows ? r : rows - 1), k);
        }
        if (base == 0 && t == 0) nan_kk = __builtin_isnan(v[0].x[0]) ? 1 : 0;
#pragma unroll
        for (int i = 0; i < LP_CHUNK; ++i) {
            const int r = base + t + LP_THREADS * i;
            const mf::V<N> c = mf::vabs(v[i]);
            if (r < rows && mf::gt(c, best)) { best = c; bi = r; }
        }
    }
    if (info != 0) return;                         // uniform
    lp_wave_argmax<N>(best, bi);
    if (lane == 0) {
#pragma unroll
        for (int l = 0; l < N; ++l) wv[wave][l] = best.x[l];
        wi[wave] = bi;
    }
    __syncthreads();
    mf::V<N> b;
#pragma unroll
    for (int l = 0; l < N; ++l) b.x[l] = wv[0][l];
    int ib = wi[0];
#pragma unroll
    for (int w = 1; w < LP_THREADS / 64; ++w) {
        mf::V<N> o;
#pragma unroll
        for (int l = 0; l < N; ++l) o.x[l] = wv[w][l];
        const int io = wi[w];
        if (mf::gt(o, b) || (mf::eq(o, b) && io < ib)) { b = o; ib = io; }
    }
    const int64_t piv = (nan_kk || ib == INT32_MAX) ? k : k + ib;
    if (t == 0) { ipiv[k] = piv + 1; st->aux = piv; }
    if (piv != k) {
        for (int64_t e = t; e < (c1 - c0) * N; e += LP_THREADS) {
            const int64_t j = c0 + e / N;
            const int l = (int)(e % N);
            double *pk = A.p + k * A.rs + j * A.cs + l * A.ls, *pp = A.p + piv * A.rs + j * A.cs + l * A.ls;
            const double x = *pk, y = *pp;
            *pk = y;
            *pp = x;
        }
    }
}

// The panel's interchanges ipiv[first .. first + count) applied in order to
// the columns [a0, a0 + na) and [b0, b0 + nb), one thread per (column, limb).
// No status check: after a failing pivot MFLA has already swapped whole rows
// up to that column, and the unused ipiv entries are the identity.
template <int N>
__global__ void k_lu_laswp(DV A, int64_t first, int64_t count, const int64_t *ipiv, int64_t a0, int64_t na, int64_t b0, int64_t nb) {
    __builtin_amdgcn_s_setprio(3);
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (na + nb) * N) return;
    const int64_t c = e / N;
    const int l = (int)(e % N);
    const int64_t j = c < na ? a0 + c : b0 + (c - na);
    double *col = A.p + j * A.cs + l * A.ls;
    for (int64_t k = first; k < first + count; ++k) {
        const int64_t p = ipiv[k] - 1;
        if (p != k) {
            const double x = col[k * A.rs], y = col[p * A.rs];
            col[k * A.rs] = y;
            col[p * A.rs] = x;
        }
    }
}

// lu.jl:104-109 + rank-1 update lu.jl:47-82 restricted to the panel columns.
// 16 lanes per row: each lane forms the row's multiplier l = a(r, k) *
// inv(a(k, k)) and applies the msubs of columns k + 1 + (lane % 16) + 16 q
// (one for the default 16-column panels); lane 0 of
// the row stores l (the other lanes of its wave loaded a(r, k) in the same
// instruction, before that store).
constexpr int LS_LANES = 16;

template <int N>
__global__ __launch_bounds__(256, 8) void k_lu_scale_update(DV A, int64_t k, int64_t last, int64_t m, Status *st) {
    __builtin_amdgcn_s_setprio(3);
    if (st->info != 0) return;
    const mf::V<N> akk = ld<N>(A, k, k);
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (mf::is_zero(akk)) {
        if (gid == 0) st->info = k + 1;
        return;
    }
    const int64_t r = k + 1 + gid / LS_LANES;
    const int j = (int)(gid % LS_LANES);
    if (r >= m) return;
    const mf::V<N> inv = mf::inv(akk);
    const mf::V<N> l = mf::mul(ld<N>(A, r, k), inv);
    for (int64_t c = k + 1 + j; c <= last; c += LS_LANES) stv<N>(A, r, c, mf::msub(ld<N>(A, r, c), l, ld<N>(A, k, c)));
    if (j == 0) stv<N>(A, r, k, l);
}

// Sequential row interchanges on right-hand sides (solve.jl:85-113).
template <int N>
__global__ void k_apply_ipiv(DV B, int64_t count, int64_t nrhs, const int64_t *ipiv) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nrhs) return;
    for (int64_t k = 0; k < count; ++k) {
        const int64_t p = ipiv[k] - 1;
        if (p != k) swap_el<N>(B, k, c, p, c);
    }
}

template <int N>
__global__ __launch_bounds__(512) void k_ldlt_unblocked(DV A, int64_t n, DV dsub, int64_t *pivots, uint8_t *blocks, S<N> alpha,
                                                        Status *st) {
    if (st->info != 0) return;
    __shared__ double sv_buf[512 * N];
    __shared__ int64_t si_buf[512];
    __shared__ double cbuf[2 * 512 * N];
    __shared__ int64_t info_s;
    if (threadIdx.x == 0) info_s = 0;
    __syncthreads();
    ldlt_unblocked_body<N>(A, n, dsub, pivots, blocks, sv(alpha), &info_s, sv_buf, si_buf, n <= 512 ? cbuf : nullptr);
    __syncthreads();
    if (threadIdx.x == 0 && info_s != 0) st->info = info_s;
}

// Lazy blocked panel (ldlt.jl:414-593). Writes info to st->info and the
// (possibly extended) last panel column to st->aux. With `spec` (the host
// launched the following panels assuming this one ends at reqlast), a panel
// that extends by a 2x2 pivot records kPanelExtended instead of 0, so every
// kernel launched after it returns at once and the host resumes after it.
//
// A Schur entry (i1, i2) -> (row, col) = (max, min) is taken against the
// pivots q = pfirst..k-1 of this panel in MFLA's order: a 1x1 pivot
// subtracts A(row, q) * (A(q, q) * A(col, q)), a 2x2 pivot subtracts
// A(row, q) * c1 + A(row, q + 1) * c2 with c1 = A(q, q) * A(col, q) +
// dsub(q) * A(col, q + 1), c2 = dsub(q) * A(col, q) + A(q + 1, q + 1) *
// A(col, q + 1).
//
// When the panel (with a possible 2x2 extension) is at most W = the default
// width + 1 columns wide, every evaluation issues all of its panel loads at
// once (clamped indices, no branch on a loaded value) and takes the pivot
// kinds, the pivots' diagonals and dsub from LDS; a column pass (fixed col)
// takes the row-independent factors (the parenthesized products above) from
// LDS, and the row pass (fixed row) takes A(row, q) from LDS. Each value is
// formed by the same operations as the per-element form, which wider panels
// use.
template <int N> constexpr int ldlt_pw() { return N <= 2 ? 17 : (N == 3 ? 13 : 9); }

template <int N>
struct LazyPanelEntries {
    static constexpr int W = ldlt_pw<N>();
    DV A, dsub;
    const uint8_t *blocks;
    int64_t pfirst, kk;
    bool hoist;
    int *kind;          // [W]: blocks[q] (1: 1x1 pivot, 2: first column of a 2x2, 0: its second)
    double *dqq, *dsq;  // [W * N]: A(q, q), dsub(q)
    double *coef;       // [2][W * N]: column factors of prep(col, set)
    double *rowv;       // [W * N]: A(row, q) of prow(row)
    double *er, *ec;    // [W * N]: A(row, q), A(col, q) of one uniform entry()
    double *keepbuf;    // device [2 * N * n]: the column pass's values, limb l of row r at l * n + r
    int64_t n;
    mf::V<N> akk;       // the last entry(kk, kk) (row kk keeps it until a pivot moves row kk)
    // entry(k, k) beside the first column-pass row: a gain where registers
    // allow (x1 / x2), a loss for x4
    static constexpr bool kFuseDiag = N <= 2;
    bool rowk;          // er and ec hold A(kk, q) (staged by setup, until an interchange moves row kk or entry() restages)

    // kk is the same in every lane: read it as a scalar so the per-pivot
    // branches below are scalar branches
    __device__ int nt() const { return __builtin_amdgcn_readfirstlane((int)(kk - pfirst)); }
    __device__ int kind_of(int j) const { return __builtin_amdgcn_readfirstlane(kind[j]); }

    // keeps the LDS reads of one pivot's term from being hoisted above the
    // previous term (all of them live at once would spill)
    __device__ static void lds_fence() { asm volatile("" ::: "memory"); }

    // the per-element form
    __device__ mf::V<N> slow(int64_t row, int64_t col) const {
        mf::V<N> v = ld<N>(A, row, col);
        int64_t q = pfirst;
        while (q < kk) {
            if (blocks[q] == 1) {
                v = mf::sub(v, mf::mul(ld<N>(A, row, q), mf::mul(ld<N>(A, q, q), ld<N>(A, col, q))));
                q += 1;
            } else {
                const mf::V<N> rf = ld<N>(A, row, q), rs = ld<N>(A, row, q + 1), cf = ld<N>(A, col, q), cs = ld<N>(A, col, q + 1);
                const mf::V<N> dq = ld<N>(dsub, q, 0);
                const mf::V<N> c1 = mf::add(mf::mul(ld<N>(A, q, q), cf), mf::mul(dq, cs));
                const mf::V<N> c2 = mf::add(mf::mul(dq, cf), mf::mul(ld<N>(A, q + 1, q + 1), cs));
                v = mf::sub(v, mf::add(mf::mul(rf, c1), mf::mul(rs, c2)));
                q += 2;
            }
        }
        return v;
    }

    // the panel columns of one row, all loads issued before any use
    __device__ void load_row(int64_t r, mf::V<N> (&x)[W]) const {
        const int last = nt() > 0 ? nt() - 1 : 0;
#pragma unroll
        for (int j = 0; j < W; ++j) x[j] = ld<N>(A, r, pfirst + (j < last ? j : last));
    }

    // v minus the pivots' terms; fx(j), fy(j) give A(row, q), A(col, q)
    template <typename FX, typename FY>
    __device__ mf::V<N> reduce(mf::V<N> v, FX fx, FY fy) const {
        const int m = nt();
#pragma unroll
        for (int j = 0; j < W; ++j) {
            lds_fence();
            if (j < m) {
                const int t = kind_of(j);
                if (t == 1) {
                    v = mf::sub(v, mf::mul(fx(j), mf::mul(lds_get<N>(&dqq[j * N]), fy(j))));
                } else if (t == 2 && j + 1 < W) {
                    const mf::V<N> dq = lds_get<N>(&dsq[j * N]), cf = fy(j), cs = fy(j + 1);
                    const mf::V<N> c1 = mf::add(mf::mul(lds_get<N>(&dqq[j * N]), cf), mf::mul(dq, cs));
                    const mf::V<N> c2 = mf::add(mf::mul(dq, cf), mf::mul(lds_get<N>(&dqq[(j + 1) * N]), cs));
                    v = mf::sub(v, mf::add(mf::mul(fx(j), c1), mf::mul(fx(j + 1), c2)));
                }
            }
        }
        return v;
    }

    // v minus the pivots' terms of a prepared column: x = A(row, q) in registers
    __device__ mf::V<N> reduce_coef(mf::V<N> v, const mf::V<N> (&x)[W], int set) const {
        const int m = nt();
#pragma unroll
        for (int j = 0; j < W; ++j) {
            lds_fence();
            if (j < m) {
                const int t = kind_of(j);
                const double *c = &coef[(set * W + j) * N];
                if (t == 1) v = mf::sub(v, mf::mul(x[j], lds_get<N>(c)));
                else if (t == 2 && j + 1 < W) v = mf::sub(v, mf::add(mf::mul(x[j], lds_get<N>(c)), mf::mul(x[j + 1], lds_get<N>(c + N))));
            }
        }
        return v;
    }

    // kinds and diagonals of the pivots before kk (every thread, once per step)
    __device__ void setup() {
        __syncthreads();
        if (hoist && (int)threadIdx.x < nt()) {
            const int j = threadIdx.x;
            const int64_t q = pfirst + j;
            kind[j] = blocks[q];
            lds_put<N>(&dqq[j * N], ld<N>(A, q, q));
            lds_put<N>(&dsq[j * N], ld<N>(dsub, q, 0));
            const mf::V<N> a = ld<N>(A, kk, q);
            lds_put<N>(&er[j * N], a);
            lds_put<N>(&ec[j * N], a);
        }
        rowk = true;
        __syncthreads();
    }

    // one entry, the same for every thread (called by every thread)
    __device__ mf::V<N> entry(int64_t i1, int64_t i2) {
        const int64_t row = i1 > i2 ? i1 : i2, col = i1 > i2 ? i2 : i1;
        if (!hoist) {
            const mf::V<N> v = slow(row, col);
            if (row == kk && col == kk) akk = v;
            return v;
        }
        if (!(rowk && row == kk && col == kk)) {
            __syncthreads();                          // earlier readers of er / ec are done
            if ((int)threadIdx.x < nt()) {
                const int j = threadIdx.x;
                lds_put<N>(&er[j * N], ld<N>(A, row, pfirst + j));
                lds_put<N>(&ec[j * N], ld<N>(A, col, pfirst + j));
            }
            rowk = false;
            __syncthreads();
        }
        const mf::V<N> v = reduce(ld<N>(A, row, col), [&](int j) { return lds_get<N>(&er[j * N]); }, [&](int j) { return lds_get<N>(&ec[j * N]); });
        if (row == kk && col == kk) akk = v;
        return v;
    }

    __device__ void prep(int64_t col, int set) {
        __syncthreads();                              // earlier readers of coef[set] (and of A) are done
        if (hoist && (int)threadIdx.x < nt()) {
            const int j = threadIdx.x;
            const int64_t q = pfirst + j;
            double *c = &coef[(set * W + j) * N];
            const int t = kind[j];
            const bool staged = rowk && col == kk;    // A(kk, q) in er
            if (t == 1) {
                lds_put<N>(c, mf::mul(lds_get<N>(&dqq[j * N]), staged ? lds_get<N>(&er[j * N]) : ld<N>(A, col, q)));
            } else if (t == 2) {
                const mf::V<N> cf = staged ? lds_get<N>(&er[j * N]) : ld<N>(A, col, q);
                const mf::V<N> cs = staged ? lds_get<N>(&er[(j + 1) * N]) : ld<N>(A, col, q + 1);
                const mf::V<N> dq = lds_get<N>(&dsq[j * N]);
                lds_put<N>(c, mf::add(mf::mul(lds_get<N>(&dqq[j * N]), cf), mf::mul(dq, cs)));
                lds_put<N>(c + N, mf::add(mf::mul(dq, cf), mf::mul(lds_get<N>(&dqq[(j + 1) * N]), cs)));
            }
        }
        __syncthreads();
    }
    __device__ void prep(int64_t col) { prep(col, 0); }

    // entry(r, col) for r > col after prep(col, set)
    __device__ mf::V<N> centry(int64_t r, int64_t col, int set = 0) const {
        if (!hoist) return slow(r, col);
        mf::V<N> x[W];
        load_row(r, x);
        return reduce_coef(ld<N>(A, r, col), x, set);
    }
    // entry(kk, kk) at the start of a step (row kk staged in er / ec by
    // setup): no barrier, any subset of threads
    __device__ mf::V<N> diag_entry(int64_t k) {
        if (!hoist) { akk = slow(k, k); return akk; }
        akk = reduce(ld<N>(A, k, k), [&](int j) { return lds_get<N>(&er[j * N]); }, [&](int j) { return lds_get<N>(&ec[j * N]); });
        return akk;
    }

    // centry(r, k, 0) after prep(k, 0) and entry(k, k), their chains in one
    // loop (same operations as centry and entry each)
    __device__ void centry_diag(int64_t r, int64_t k, mf::V<N> &v, mf::V<N> &d) {
        if (!hoist) { v = slow(r, k); d = slow(k, k); akk = d; return; }
        mf::V<N> x[W];
        load_row(r, x);
        v = ld<N>(A, r, k);
        d = ld<N>(A, k, k);
        const int m = nt();
#pragma unroll
        for (int j = 0; j < W; ++j) {
            lds_fence();
            if (j < m) {
                const int t = kind_of(j);
                const double *c = &coef[j * N];
                if (t == 1) {
                    v = mf::sub(v, mf::mul(x[j], lds_get<N>(c)));
                    d = mf::sub(d, mf::mul(lds_get<N>(&er[j * N]), mf::mul(lds_get<N>(&dqq[j * N]), lds_get<N>(&ec[j * N]))));
                } else if (t == 2 && j + 1 < W) {
                    v = mf::sub(v, mf::add(mf::mul(x[j], lds_get<N>(c)), mf::mul(x[j + 1], lds_get<N>(c + N))));
                    const mf::V<N> dq = lds_get<N>(&dsq[j * N]), cf = lds_get<N>(&ec[j * N]), cs = lds_get<N>(&ec[(j + 1) * N]);
                    const mf::V<N> c1 = mf::add(mf::mul(lds_get<N>(&dqq[j * N]), cf), mf::mul(dq, cs));
                    const mf::V<N> c2 = mf::add(mf::mul(dq, cf), mf::mul(lds_get<N>(&dqq[(j + 1) * N]), cs));
                    d = mf::sub(d, mf::add(mf::mul(lds_get<N>(&er[j * N]), c1), mf::mul(lds_get<N>(&er[(j + 1) * N]), c2)));
                }
            }
        }
        akk = d;
    }

    // column-pass values (set 0), read back when the pivot leaves column k
    // in place and, after an interchange with k + 1, by the 2x2 step
    __device__ void keep(int64_t r, const mf::V<N> &v) const {
#pragma unroll
        for (int l = 0; l < N; ++l) keepbuf[l * n + r] = v.x[l];
    }
    // centry(r, imax) of the row-max pass, r > imax, into set 1: after the
    // interchange of imax with k (1x1) or k + 1 (2x2) it is the new
    // entry(r, k) or entry(r, k + 1) of rows below imax
    __device__ void keep_row(int64_t r, const mf::V<N> &v) const {
#pragma unroll
        for (int l = 0; l < N; ++l) keepbuf[(N + l) * n + r] = v.x[l];
    }
    __device__ mf::V<N> kept(int64_t r, int set = 0) const {
        mf::V<N> v;
#pragma unroll
        for (int l = 0; l < N; ++l) v.x[l] = keepbuf[(set * N + l) * n + r];
        return v;
    }
    __device__ void keep2(int64_t r, const mf::V<N> &a, const mf::V<N> &b) const {
#pragma unroll
        for (int l = 0; l < N; ++l) {
            keepbuf[l * n + r] = a.x[l];
            keepbuf[(N + l) * n + r] = b.x[l];
        }
    }

    __device__ void prow(int64_t row) {
        __syncthreads();
        if (hoist && (int)threadIdx.x < nt()) lds_put<N>(&rowv[threadIdx.x * N], ld<N>(A, row, pfirst + threadIdx.x));
        __syncthreads();
    }

    // entry(row, c) for c < row after prow(row)
    __device__ mf::V<N> rentry(int64_t row, int64_t c) const {
        if (!hoist) return slow(row, c);
        mf::V<N> y[W];
        load_row(c, y);
        return reduce(ld<N>(A, row, c), [&](int j) { return lds_get<N>(&rowv[j * N]); }, [&](int j) { return y[j]; });
    }
};

template <int N>
__global__ __launch_bounds__(512) void k_ldlt_panel(DV A, int64_t n, int64_t pfirst, int64_t reqlast, DV dsub, int64_t *pivots,
                                                    uint8_t *blocks, S<N> alpha_s, double *keepbuf, int spec, Status *st) {
    if (st->info != 0) return;
    constexpr int W = ldlt_pw<N>();
    __shared__ double sv_buf[512 * N];
    __shared__ int64_t si_buf[512];
    __shared__ int kind[W];
    __shared__ double dqq[W * N], dsq[W * N], coef[2 * W * N], rowv[W * N], er[W * N], ec[W * N];
    const mf::V<N> alpha = sv(alpha_s);
    LazyPanelEntries<N> e{A, dsub, blocks, pfirst, pfirst, reqlast - pfirst + 2 <= W, kind, dqq, dsq, coef, rowv, er, ec, keepbuf, n};
    int64_t plast = reqlast;
    int64_t k = pfirst;
    int64_t info = 0, frozen = 0;
    // the upper triangle of row k: the symmetric block's row as it stands
    // after the step's interchange (MFLA's upper row k is final from then on)
    auto freeze_row = [&](int64_t row, int64_t from) {
        for (int64_t r = from + threadIdx.x; r < n; r += blockDim.x) stv<N>(A, row, r, ld<N>(A, r, row));
    };
    while (k <= plast) {
        e.kk = k;
        e.setup();
        BKChoice ch = wg_select_bk<N>(n, k, alpha, e, sv_buf, si_buf);
        if (ch.bs == 0) { freeze_row(k, k + 1); info = k + 1; frozen = k + 1; break; }
        if (ch.bs == 2 && k == plast) {
            if (!(k < n - 1)) { freeze_row(k, k + 1); info = k + 1; frozen = k + 1; break; }
            plast += 1;
        }
        if (ch.bs == 1) {
            if (threadIdx.x == 0) { pivots[k] = ch.pivot + 1; blocks[k] = 1; }
            __syncthreads();
            wg_sym_swap_lower<N>(A, n, k, ch.pivot);
            if (ch.pivot != k) e.rowk = false;       // row k now holds the pivot row
            freeze_row(k, k + 1);                    // each thread's rows (it stores them below)
            const mf::V<N> d = ch.pivot == k ? e.akk : e.entry(k, k);
            frozen = k + 1;
            if (mf::is_zero(d)) { info = k + 1; break; }
            if (ch.pivot == k) {
                // no interchange: column k is what the column pass saw (each
                // thread reads back its own rows)
                __syncthreads();
                if (threadIdx.x == 0) stv<N>(A, k, k, d);
                for (int64_t r = k + 1 + threadIdx.x; r < n; r += blockDim.x) stv<N>(A, r, k, mf::div(e.kept(r), d));
            } else {
                // rows below the pivot row: the row-max pass formed their
                // new entry (set 1); the others are formed here
                e.prep(k, 0);
                if (threadIdx.x == 0) stv<N>(A, k, k, d);
                for (int64_t r = k + 1 + threadIdx.x; r < n; r += blockDim.x)
                    stv<N>(A, r, k, mf::div(r > ch.pivot ? e.kept(r, 1) : e.centry(r, k, 0), d));
            }
            __syncthreads();
            k += 1;
        } else {
            if (threadIdx.x == 0) { pivots[k] = ch.pivot + 1; blocks[k] = 2; blocks[k + 1] = 0; }
            __syncthreads();
            wg_sym_swap_lower<N>(A, n, k + 1, ch.pivot);
            freeze_row(k, k + 2);
            freeze_row(k + 1, k + 2);
            if (threadIdx.x == 0) stv<N>(A, k, k + 1, ld<N>(A, k + 1, k));
            frozen = k + 2;
            // the interchange of k + 1 and the pivot leaves row k, so d11 is
            // the entry the pivot choice formed
            const mf::V<N> d11 = e.akk, d21 = e.entry(k + 1, k), d22 = e.entry(k + 1, k + 1);
            e.prep(k + 1, 1);
            if (threadIdx.x == 0) { stv<N>(A, k, k, d11); stv<N>(A, k + 1, k + 1, d22); stv<N>(dsub, k, 0, d21); }
            __syncthreads();
            mf::V<N> s1, s2;
            if (!solve2x2<N>(d11, d21, d22, mf::zero<N>(), mf::zero<N>(), s1, s2)) { info = k + 1; break; }
            int64_t myfail = INT64_MAX;
            // the interchange of k + 1 and the pivot row p leaves column k:
            // entry(r, k) is the column pass's value of row r (of row k + 1
            // for r = p); entry(r, k + 1) of a row below p is the row-max
            // pass's value (set 1), the others are formed here. Each thread
            // reads only its own rows' slots (and slot k + 1) before writing
            // its rows' solutions into them.
            const int64_t p = ch.pivot;
            for (int64_t r = k + 2 + threadIdx.x; r < n; r += blockDim.x) {
                const mf::V<N> f = e.kept(r == p ? k + 1 : r, 0);
                const mf::V<N> s = r > p ? e.kept(r, 1) : e.centry(r, k + 1, 1);
                bool ok = solve2x2<N>(d11, d21, d22, f, s, s1, s2);
                if (!(ok && mf::is_finite(s1) && mf::is_finite(s2)) && r < myfail) myfail = r;
                e.keep2(r, s1, s2);                   // read back below by this thread (rows before the first failure)
            }
            const int64_t firstfail = wg_min_i64(myfail, si_buf);
            for (int64_t r = k + 2 + threadIdx.x; r < n && r < firstfail; r += blockDim.x) {
                stv<N>(A, r, k, e.kept(r, 0));
                stv<N>(A, r, k + 1, e.kept(r, 1));
            }
            __syncthreads();
            if (firstfail != INT64_MAX) { info = k + 1; break; }
            if (threadIdx.x == 0) { stv<N>(A, k + 1, k, mf::zero<N>()); stv<N>(A, k, k + 1, mf::zero<N>()); }
            __syncthreads();
            k += 2;
        }
    }
    if (threadIdx.x == 0) {
        st->info = (info == 0 && spec && plast != reqlast) ? panel_stop_code(pfirst, false) : info;
        st->aux = info > 0 ? frozen : plast;
    }
}

// ---------------------------------------------------------------------------
// The lazy panel over G co-resident workgroups (one launch per panel). Every per-element operation is that of k_ldlt_panel;
// the work is spread and the serial chain of a step kept short:
//  * row threads (the first GP_T of a workgroup): thread g of the grid owns
//    the rows pfirst + 1 + g + m * (G * GP_T) for the whole panel; the panel
//    entries of its first row stay in registers between steps (reloaded after
//    an interchange) and that row's next original entry is loaded while the
//    grid barrier is pending;
//  * a helper wave (the last 64 threads) forms entry(k, k) beside the column
//    pass, and after the barrier the next pivot row's new entry;
//  * between a workgroup's arrival at the barrier and its wait, the row
//    threads divide their column values by entry(k, k): a step that takes a
//    1x1 pivot without interchange (the usual case) then only stores (and the
//    helper divides one value) after the barrier.
// Per step: column pass and each workgroup's first strict maximum; barrier;
// every workgroup reduces the partials in the same order and takes the same
// Bunch-Kaufman decision (the first strict maximum is a total order on
// (value, row), so any reduction order gives the same winner); the row-max
// pass likewise when needed (one barrier); an interchange -- the row part,
// the column part and the four crossing elements (composed as the row swap
// followed by the column swap) in one pass -- and a barrier; the step's
// columns on the owned rows. A step with an interchange ends with a barrier.
// One without needs none: every value a workgroup reads in the next step
// before its barrier is its own, was written before this step's barrier, or
// is the next pivot row's new entry, which each workgroup forms and stores
// itself (the same bits as the owner's) from the owner's column value,
// published before this step's barrier. (Another row's newest entry is
// written after that barrier: it is read only after the next one.)
//
// The barrier counts arrivals on st->sync (monotonic within a call;
// st->sync_base holds its value when a launch starts); a wait is bounded: a
// workgroup that waits ~1 s records kGridTimeout in st->info and leaves, and
// the others then time out at their next barrier.
constexpr int GP_T = 192;                 // row threads per workgroup
constexpr int GP_THREADS = GP_T + 64;     // and the helper wave (one wave per SIMD)
constexpr int GP_MAXG = 64;               // workgroups per launch (co-resident)
constexpr int64_t kGridTimeout = -2;      // st->info after a barrier timed out

__host__ __device__ constexpr size_t ldlt_grid_partials(int N) { return (size_t)GP_MAXG * (2 * (N + 1) + N + 1); }

struct GridBarrier {
    unsigned long long *ctr;
    unsigned long long target, G;
    int *ok;                              // LDS
    // every thread: this workgroup's writes so far reach the device scope
    __device__ void arrive() {
        if (G > 1) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (G == 1) return;
        target += G;
        if (threadIdx.x == 0) __hip_atomic_fetch_add(ctr, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // every thread: the other workgroups' writes before their arrival are seen
    __device__ bool wait() {
        if (G == 1) { __syncthreads(); return true; }
        if (threadIdx.x == 0) {
            int good = 1;
            unsigned spins = 0;
            while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                __builtin_amdgcn_s_sleep(2);
                if (++spins > (1u << 20)) { good = 0; break; }
            }
            *ok = good;
        }
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        return *ok != 0;
    }
    __device__ bool sync() { arrive(); return wait(); }
};

// workgroup max / min over the waves (butterfly in each wave, then the waves
// in order); for finite values the result does not depend on the order
template <int N>
__device__ mf::V<N> waves_max(mf::V<N> v, double *sv_buf) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        mf::V<N> b;
#pragma unroll
        for (int l = 0; l < N; ++l) b.x[l] = __shfl_xor(v.x[l], off, 64);
        v = mf::vmax(v, b);
    }
    if (lane == 0) lds_put<N>(&sv_buf[wave * N], v);
    __syncthreads();
    v = lds_get<N>(&sv_buf[0]);
    for (int w = 1; w < nw; ++w) v = mf::vmax(v, lds_get<N>(&sv_buf[w * N]));
    __syncthreads();
    return v;
}
__device__ int64_t waves_min(int64_t v, int64_t *si_buf) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const int64_t b = __shfl_xor(v, off, 64);
        v = b < v ? b : v;
    }
    if (lane == 0) si_buf[wave] = v;
    __syncthreads();
    v = si_buf[0];
    for (int w = 1; w < nw; ++w) v = si_buf[w] < v ? si_buf[w] : v;
    __syncthreads();
    return v;
}

template <int N>
__global__ __launch_bounds__(GP_THREADS) void k_ldlt_panel_grid(DV A, int64_t n, int64_t pfirst, int64_t reqlast, DV dsub, int64_t *pivots,
                                                               uint8_t *blocks, S<N> alpha_s, double *keepbuf, double *part, int spec,
                                                               int trunc, int64_t wld, Status *st) {
    if (st->info != 0) return;
    constexpr int W = ldlt_pw<N>();
    constexpr int NW = GP_THREADS / 64;
    __shared__ double sv_buf[NW * N];
    __shared__ int64_t si_buf[NW];
    __shared__ int kind[W];
    __shared__ double dqq[W * N], dsq[W * N], coef[2 * W * N], rowv[W * N], er[W * N], ec[W * N];
    __shared__ double s_akk[N], s_dpp[N];
    __shared__ int bar_ok;
    const mf::V<N> alpha = sv(alpha_s);
    const int G = (int)gridDim.x, wg = (int)blockIdx.x;
    const bool helper = threadIdx.x >= GP_T;
    const int64_t GT = (int64_t)G * GP_T;
    const int64_t gt = (int64_t)wg * GP_T + (helper ? 0 : threadIdx.x);
    const int64_t gstart = helper ? n : gt;         // this thread's first index of a grid-strided pass
    GridBarrier bar{&st->sync, st->sync_base, (unsigned long long)G, &bar_ok};
    LazyPanelEntries<N> e{A, dsub, blocks, pfirst, pfirst, reqlast - pfirst + 2 <= W, kind, dqq, dsq, coef, rowv, er, ec, keepbuf, n};
    // partials: column maxima (two sets, by step parity), row maxima, first failures
    double *p_col = part, *p_row = part + 2 * GP_MAXG * (N + 1);
    int64_t *p_fail = reinterpret_cast<int64_t *>(p_row + GP_MAXG * N);
    // quotients and 2x2 solutions of the owned rows wait in sets 2 / 3 of the keep buffer
    auto sol = [&](int64_t r, int set) {
        mf::V<N> v;
        for (int l = 0; l < N; ++l) v.x[l] = keepbuf[((2 + set) * N + l) * n + r];
        return v;
    };
    auto put_sol = [&](int64_t r, int set, const mf::V<N> &a) {
        for (int l = 0; l < N; ++l) keepbuf[((2 + set) * N + l) * n + r] = a.x[l];
    };
    // the first owned row (registers) and the further ones
    const int64_t r0 = helper ? n : pfirst + 1 + gt;
    auto more_rows = [&](int64_t after) {           // first owned row > after, beyond r0
        if (helper) return n;
        int64_t r = r0 + GT;
        if (r <= after) r += ((after - r) / GT + 1) * GT;
        return r;
    };
    auto first_row = [&](int64_t after) { return (r0 > after) ? r0 : more_rows(after); };
    auto better = [](const mf::V<N> &b, int64_t ib, const mf::V<N> &a, int64_t ia) {
        if (ib < 0) return false;
        if (ia < 0) return true;
        if (mf::gt(b, a)) return true;
        return mf::eq(a, b) && ib < ia;
    };
    const bool regs = e.hoist && r0 < n;             // row r0's finalized panel entries live in xr
    mf::V<N> xr[W];
    bool xr_ok = false;                              // xr holds A(r0, pfirst + j), j < nt
    mf::V<N> aorig;                                  // A(r0, k) when aorig_ok
    bool aorig_ok = false;
    int64_t plast = reqlast, k = pfirst, info = 0;
    int64_t frozen = 0;                              // after a failure: rows before it have their upper row written
    int64_t tstop = -1;                              // trunc: the column that needed the row-max pass
    bool ok = true;
    unsigned it = 0;
    while (k <= plast) {
        e.kk = k;
        e.setup();
        const int nt = (int)(k - pfirst);
        // ---------------------------------------------- pivot choice
        int bs;
        int64_t piv = k;
        mf::V<N> absakk;
        mf::V<N> v0 = mf::zero<N>(), q0 = mf::zero<N>();
        mf::V<N> dpp = mf::zero<N>();                // entry(imax, imax) when the row-max pass ran
        mf::V<N> a_kk = mf::zero<N>();               // helper, trunc: A(k, k) before this step
        if (k == n - 1) {
            absakk = mf::vabs(e.entry(k, k));
            bs = mf::is_zero(absakk) ? 0 : 1;
        } else {
            e.prep(k);
            mf::V<N> best = mf::zero<N>();
            int64_t bi = -1;
            if (!helper) {
                if (r0 > k && r0 < n) {
                    if (regs) {
                        if (!xr_ok) {
                            const int last = nt > 0 ? nt - 1 : 0;
#pragma unroll
                            for (int j = 0; j < W; ++j) xr[j] = ld<N>(A, r0, pfirst + (j < last ? j : last));
                            xr_ok = true;
                        }
                        if (!aorig_ok) aorig = ld<N>(A, r0, k);
                        stv<N>(A, k, r0, aorig);
                        v0 = e.reduce_coef(aorig, xr, 0);
                    } else {
                        stv<N>(A, k, r0, ld<N>(A, r0, k));
                        v0 = e.centry(r0, k);
                    }
                    e.keep(r0, v0);
                    best = mf::vabs(v0);
                    bi = r0;
                }
                for (int64_t r = more_rows(k); r < n; r += GT) {
                    stv<N>(A, k, r, ld<N>(A, r, k));
                    const mf::V<N> v = e.centry(r, k);
                    e.keep(r, v);
                    const mf::V<N> c = mf::vabs(v);
                    if (bi < 0 || mf::gt(c, best)) { best = c; bi = r; }
                }
            } else {
                // entry(k, k) beside the column pass (row k's entries: its
                // newest one is this workgroup's own store of the previous
                // step). With row k staged, coef(j) = dqq(j) * A(k, q) is the
                // inner product of entry()'s terms, so the chain takes row k
                // from registers and the prepared factors like the row threads'
                mf::V<N> d0;
                if (trunc) a_kk = ld<N>(A, k, k);
                if (e.hoist) {
                    mf::V<N> xk[W];
#pragma unroll
                    for (int j = 0; j < W; ++j) xk[j] = lds_get<N>(&er[j * N]);
                    d0 = e.reduce_coef(ld<N>(A, k, k), xk, 0);
                } else {
                    d0 = e.diag_entry(k);
                }
                if (threadIdx.x == GP_T) lds_put<N>(s_akk, d0);
            }
            wg_argmax_lead<N>(best, bi, sv_buf, si_buf);
            double *pc = p_col + (it & 1) * GP_MAXG * (N + 1);
            if (threadIdx.x == 0) {
                for (int l = 0; l < N; ++l) pc[wg * (N + 1) + l] = best.x[l];
                pc[wg * (N + 1) + N] = __builtin_bit_cast(double, bi);
            }
            bar.arrive();
            e.akk = lds_get<N>(s_akk);
            absakk = mf::vabs(e.akk);
            // while the barrier is pending: the quotients of a 1x1 pivot k
            // without interchange, and the next original entry of row r0
            if (!helper) {
                if (r0 > k && r0 < n) q0 = mf::div(v0, e.akk);
                for (int64_t r = more_rows(k); r < n; r += GT) put_sol(r, 0, mf::div(e.kept(r), e.akk));
                aorig_ok = false;
                if (regs && r0 > k + 1) { aorig = ld<N>(A, r0, k + 1); aorig_ok = true; }
            }
            if (!(ok = bar.wait())) break;
            mf::V<N> colmax = mf::zero<N>();
            int64_t imax = -1;
            for (int w = 0; w < G; ++w) {
                mf::V<N> b;
                for (int l = 0; l < N; ++l) b.x[l] = pc[w * (N + 1) + l];
                const int64_t ib = __builtin_bit_cast(int64_t, pc[w * (N + 1) + N]);
                if (better(b, ib, colmax, imax)) { colmax = b; imax = ib; }
            }
            if (mf::eq(mf::vmax(absakk, colmax), mf::zero<N>())) {
                bs = 0;
            } else if (mf::ge(absakk, mf::mul(alpha, colmax))) {
                bs = 1;
            } else if (trunc) {
                // the row-max pass reads past the panel's columns, where the
                // previous panel's update may still be running: stop here
                // (every workgroup takes this branch at the same k)
                tstop = k;
                break;
            } else {
                mf::V<N> rm = mf::zero<N>();
                e.prow(imax);
                // the helper forms entry(imax, imax) from the staged row (the
                // operations of entry()) while the row threads run the pass
                if (helper) {
                    const mf::V<N> v = e.hoist ? e.reduce(ld<N>(A, imax, imax), [&](int j) { return lds_get<N>(&rowv[j * N]); },
                                                          [&](int j) { return lds_get<N>(&rowv[j * N]); })
                                               : e.slow(imax, imax);
                    if (threadIdx.x == GP_T) lds_put<N>(s_dpp, v);
                }
                for (int64_t c = k + gstart; c < imax; c += GT) rm = mf::vmax(rm, mf::vabs(e.rentry(imax, c)));
                e.prep(imax);
                for (int64_t r = first_row(imax); r < n; r += GT) {
                    const mf::V<N> v = e.centry(r, imax);
                    e.keep_row(r, v);
                    rm = mf::vmax(rm, mf::vabs(v));
                }
                rm = waves_max<N>(rm, sv_buf);
                if (threadIdx.x == 0)
                    for (int l = 0; l < N; ++l) p_row[wg * N + l] = rm.x[l];
                if (!(ok = bar.sync())) break;
                mf::V<N> rowmax = mf::zero<N>();
                for (int w = 0; w < G; ++w) {
                    mf::V<N> b;
                    for (int l = 0; l < N; ++l) b.x[l] = p_row[w * N + l];
                    rowmax = mf::vmax(rowmax, b);
                }
                dpp = lds_get<N>(s_dpp);
                if (mf::ge(absakk, mf::mul(mf::mul(alpha, colmax), mf::div(colmax, rowmax)))) {
                    bs = 1;
                } else if (mf::ge(mf::vabs(dpp), mf::mul(alpha, rowmax))) {
                    bs = 1;
                    piv = imax;
                } else {
                    bs = 2;
                    piv = imax;
                }
            }
        }
        ++it;
        if (bs == 0) { info = k + 1; frozen = k + (bs == 2 ? 2 : 1); break; }
        if (bs == 2 && k == plast) {
            if (!(k < n - 1)) { info = k + 1; frozen = k + (bs == 2 ? 2 : 1); break; }
            plast += 1;
        }
        // ---------------------------------------------- interchange
        const bool swapped = piv != (bs == 1 ? k : k + 1);
        Prepared2x2<N> pr2;
        if (swapped) {
            // the lower triangle's part of the symmetric interchange (rows f, s
            // left of f; (r, f) <-> (s, r) for f < r < s; columns f, s below s;
            // the two diagonals): disjoint pairs, one pass
            const int64_t f = bs == 1 ? k : k + 1, s = piv;
            for (int64_t c = gstart; c < f; c += GT) swap_el<N>(A, f, c, s, c);
            for (int64_t r = f + 1 + gstart; r < s; r += GT) swap_el<N>(A, r, f, s, r);
            for (int64_t r = s + 1 + gstart; r < n; r += GT) swap_el<N>(A, r, f, r, s);
            if (gt == 0 && !helper) swap_el<N>(A, f, f, s, s);
            bar.arrive();
            // the 2x2 pivot block's own part of the solves while the barrier
            // is pending: after the interchange entry(k + 1, k) is the column
            // pass's value of row piv and entry(k + 1, k + 1) is entry(piv, piv)
            if (bs == 2) pr2 = prepare2x2<N>(e.akk, e.kept(piv, 0), dpp);
            if (!(ok = bar.wait())) break;
            if (bs == 1) e.rowk = false;
        } else if (bs == 2) {
            pr2 = prepare2x2<N>(e.akk, e.kept(piv, 0), dpp);
        }
        if (threadIdx.x == 0) {
            pivots[k] = piv + 1;
            blocks[k] = (uint8_t)bs;
            if (bs == 2) blocks[k + 1] = 0;
        }
        // ---------------------------------------------- the step's columns
        if (bs == 1 && !swapped) {
            // the quotients formed while the barrier was pending
            const mf::V<N> d = e.akk;
            if (mf::is_zero(d)) { info = k + 1; frozen = k + (bs == 2 ? 2 : 1); break; }
            if (!helper) {
                if (r0 > k && r0 < n) {
                    stv<N>(A, r0, k, q0);
                    if (regs) {
#pragma unroll
                        for (int j = 0; j < W; ++j)
                            if (j == nt) xr[j] = q0;
                    }
                }
                for (int64_t r = more_rows(k); r < n; r += GT) stv<N>(A, r, k, sol(r, 0));
            }
            if (helper) {
                // the next pivot row's new entry from the owner's column value
                // (published before the barrier): the same bits as the owner's
                const mf::V<N> q1 = k + 1 < n ? mf::div(e.kept(k + 1), d) : mf::zero<N>();
                if (threadIdx.x == GP_T) {
                    // trunc: the diagonal's value before the panel, kept for
                    // the restore should a later column stop the panel (set 3
                    // is the 2x2 pivots' and unused before a row-max pass)
                    if (trunc && wg == 0) put_sol(k, 1, a_kk);
                    stv<N>(A, k, k, d);
                    if (k + 1 < n) stv<N>(A, k + 1, k, q1);
                }
            }
            __syncthreads();
            k += 1;
        } else if (bs == 1) {
            for (int64_t r = first_row(k); r < n; r += GT) stv<N>(A, k, r, ld<N>(A, r, k));
            // entry(k, k) after the interchange is entry(piv, piv) before it
            // (row k now holds row piv's entries and diagonal)
            const mf::V<N> d = dpp;
            if (mf::is_zero(d)) { info = k + 1; frozen = k + (bs == 2 ? 2 : 1); break; }
            e.prep(k, 0);
            for (int64_t r = first_row(k); r < n; r += GT)
                stv<N>(A, r, k, mf::div(r > piv ? e.kept(r, 1) : e.centry(r, k, 0), d));
            if (!(ok = bar.sync())) break;
            if (threadIdx.x == 0) stv<N>(A, k, k, d);
            xr_ok = false;
            aorig_ok = false;
            __syncthreads();
            k += 1;
        } else {
            const int64_t p = piv;
            for (int64_t r = first_row(k + 1); r < n; r += GT) {
                stv<N>(A, k, r, ld<N>(A, r, k));
                stv<N>(A, k + 1, r, ld<N>(A, r, k + 1));
            }
            if (threadIdx.x == 0) stv<N>(A, k, k + 1, ld<N>(A, k + 1, k));
            // after the interchange of k + 1 and p, entry(k + 1, k) is the
            // column pass's value of row p and entry(k + 1, k + 1) is
            // entry(p, p) (row k + 1 holds row p's entries and diagonal)
            const mf::V<N> d11 = e.akk, d21 = e.kept(p, 0), d22 = dpp;
            e.prep(k + 1, 1);
            if (!pr2.ok) {
                if (!(ok = bar.sync())) break;
                if (threadIdx.x == 0) { stv<N>(A, k, k, d11); stv<N>(A, k + 1, k + 1, d22); stv<N>(dsub, k, 0, d21); }
                info = k + 1;
                frozen = k + 2;
                break;
            }
            // entry(r, k) of a row r >= k + 2 is the column pass's value (of
            // row k + 1 for r = p), entry(r, k + 1) of a row below p the
            // row-max pass's; the solutions wait in sets 2 / 3 until the first
            // failure is known
            int64_t myfail = INT64_MAX;
            for (int64_t r = first_row(k + 1); r < n; r += GT) {
                const mf::V<N> fv = e.kept(r == p ? k + 1 : r, 0);
                const mf::V<N> sv2 = r > p ? e.kept(r, 1) : e.centry(r, k + 1, 1);
                mf::V<N> x1, x2;
                apply2x2<N>(pr2, fv, sv2, x1, x2);
                if (!(mf::is_finite(x1) && mf::is_finite(x2)) && r < myfail) myfail = r;
                put_sol(r, 0, x1);
                put_sol(r, 1, x2);
            }
            // the next pivot row's solution, formed by every workgroup's helper
            mf::V<N> n1 = mf::zero<N>(), n2 = mf::zero<N>();
            if (helper && k + 2 < n) {
                const int64_t r = k + 2;
                const mf::V<N> fv = e.kept(r == p ? k + 1 : r, 0);
                const mf::V<N> sv2 = r > p ? e.kept(r, 1) : e.centry(r, k + 1, 1);
                apply2x2<N>(pr2, fv, sv2, n1, n2);
            }
            myfail = waves_min(myfail, si_buf);
            if (threadIdx.x == 0) p_fail[wg] = myfail;
            if (!(ok = bar.sync())) break;
            int64_t firstfail = INT64_MAX;
            for (int w = 0; w < G; ++w) firstfail = p_fail[w] < firstfail ? p_fail[w] : firstfail;
            for (int64_t r = first_row(k + 1); r < n && r < firstfail; r += GT) {
                stv<N>(A, r, k, sol(r, 0));
                stv<N>(A, r, k + 1, sol(r, 1));
            }
            if (threadIdx.x == 0) { stv<N>(A, k, k, d11); stv<N>(A, k + 1, k + 1, d22); stv<N>(dsub, k, 0, d21); }
            if (firstfail != INT64_MAX) { info = k + 1; frozen = k + (bs == 2 ? 2 : 1); break; }
            if (threadIdx.x == 0) {
                stv<N>(A, k + 1, k, mf::zero<N>());
                stv<N>(A, k, k + 1, mf::zero<N>());
            }
            if (threadIdx.x == GP_T && k + 2 < n) { stv<N>(A, k + 2, k, n1); stv<N>(A, k + 2, k + 1, n2); }
            xr_ok = false;
            aorig_ok = false;
            __syncthreads();
            k += 2;
        }
    }
    // wld > 0: the weighted panel W(r - plast - 1, q - pfirst) = L21 D of the
    // rows below the panel (ldlt.jl:595-632, the operations of
    // k_ldlt_weighted), limb stride (n - plast - 1) * wld, written over the
    // keep buffer once every workgroup is past its last read of it
    if (ok && info == 0 && tstop < 0 && wld > 0 && plast < n - 1 && (ok = bar.sync())) {
        const int64_t tcount = n - plast - 1;
        const DV Wv{keepbuf, 1, tcount, tcount * wld};
        for (int64_t r = first_row(plast); r < n; r += GT) {
            const int64_t lr = r - plast - 1;
            int64_t q = pfirst;
            while (q <= plast) {
                const int64_t lc = q - pfirst;
                if (blocks[q] == 1) {
                    stv<N>(Wv, lr, lc, mf::mul(ld<N>(A, r, q), ld<N>(A, q, q)));
                    q += 1;
                } else {
                    const mf::V<N> d11 = ld<N>(A, q, q), d21 = ld<N>(dsub, q, 0), d22 = ld<N>(A, q + 1, q + 1);
                    const mf::V<N> f = ld<N>(A, r, q), s = ld<N>(A, r, q + 1);
                    stv<N>(Wv, lr, lc, mf::add(mf::mul(f, d11), mf::mul(s, d21)));
                    stv<N>(Wv, lr, lc + 1, mf::add(mf::mul(f, d21), mf::mul(s, d22)));
                    q += 2;
                }
            }
        }
    }
    if (threadIdx.x == 0) {
        if (!ok) {
            __hip_atomic_exchange(&st->info, kGridTimeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else if (wg == 0) {
            int64_t expect = 0;
            const int64_t code = tstop >= 0                                  ? panel_stop_code(pfirst, true)
                                 : (info == 0 && spec && plast != reqlast) ? panel_stop_code(pfirst, false)
                                                                           : info;
            __hip_atomic_compare_exchange_strong(&st->info, &expect, code, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            st->aux = info > 0 ? frozen : (tstop >= 0 ? tstop : plast);
            st->sync_base = bar.target;
        }
    }
}

// After a panel in truncation mode stopped at column kstop: its columns
// [pfirst, kstop) back to the values it started from. Those columns took 1x1
// pivots without interchange, and each column pass left its column's entries
// below the diagonal in row q's upper part; the diagonals wait in set 3 of
// the keep buffer (limb l of A(q, q) at (3 N + l) n + q).
template <int N>
__global__ __launch_bounds__(256) void k_ldlt_restore(DV A, int64_t n, int64_t pfirst, const double *keepbuf) {
    const int64_t q = pfirst + blockIdx.y;
    for (int64_t r = q + 1 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        stv<N>(A, r, q, ld<N>(A, q, r));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        mf::V<N> d;
#pragma unroll
        for (int l = 0; l < N; ++l) d.x[l] = keepbuf[(3 * N + l) * n + q];
        stv<N>(A, q, q, d);
    }
}

// ldlt.jl:595-632: weighted panel W(r, local) = L21 * D for the trailing rows.
template <int N>
__global__ void k_ldlt_weighted(DV A, int64_t pfirst, int64_t plast, int64_t tfirst, int64_t tcount, DV dsub, const uint8_t *blocks,
                                DV W, const Status *st) {
    if (st->info != 0) return;
    const int64_t lr = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (lr >= tcount) return;
    const int64_t row = tfirst + lr;
    int64_t q = pfirst;
    while (q <= plast) {
        const int64_t lc = q - pfirst;
        if (blocks[q] == 1) {
            stv<N>(W, lr, lc, mf::mul(ld<N>(A, row, q), ld<N>(A, q, q)));
            q += 1;
        } else {
            const mf::V<N> d11 = ld<N>(A, q, q), d21 = ld<N>(dsub, q, 0), d22 = ld<N>(A, q + 1, q + 1);
            const mf::V<N> f = ld<N>(A, row, q), s = ld<N>(A, row, q + 1);
            stv<N>(W, lr, lc, mf::add(mf::mul(f, d11), mf::mul(s, d21)));
            stv<N>(W, lr, lc + 1, mf::add(mf::mul(f, d21), mf::mul(s, d22)));
            q += 2;
        }
    }
}

// ldlt.jl:141-153: A[c, r] = A[r, c] for r > c over an n x n view. One
// workgroup per 64 x 64 tile of the lower triangle, staged through LDS so
// that both the column reads and the transposed writes run along the unit
// stride. A tile reads only lower-triangle elements (and its own diagonal
// block) and writes only its mirror image, so workgroups never overlap.
constexpr int MIR_T = 64;

template <int N>
__global__ __launch_bounds__(256) void k_mirror_lower(DV A, int64_t n, const Status *st) {
    if (st && st->info != 0) return;
    const int64_t ti = blockIdx.y, tj = blockIdx.x;   // tile row, tile column
    if (tj > ti) return;
    __shared__ double tile[MIR_T][MIR_T + 1];
    const int64_t r0 = ti * MIR_T, c0 = tj * MIR_T;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int l = 0; l < N; ++l) {
        double *p = A.p + l * A.ls;
#pragma unroll 4
        for (int i = ty; i < MIR_T; i += 4) {
            const int64_t r = r0 + tx, c = c0 + i;
            if (r < n && c < n) tile[tx][i] = p[r * A.rs + c * A.cs];
        }
        __syncthreads();
#pragma unroll 4
        for (int i = ty; i < MIR_T; i += 4) {
            const int64_t c = c0 + tx, r = r0 + i;
            if (r < n && c < n && r > c) p[c * A.rs + r * A.cs] = tile[i][tx];
        }
        __syncthreads();
    }
}

// solve.jl:207-241: LDLT interchanges on right-hand sides, one thread per column.
template <int N>
__global__ void k_ldlt_pivots(DV B, int64_t n, int64_t nrhs, const int64_t *pivots, const uint8_t *blocks, int forward, int *bad) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nrhs) return;
    if (forward) {
        int64_t k = 0;
        while (k < n) {
            if (blocks[k] == 1) { int64_t p = pivots[k] - 1; if (p != k) swap_el<N>(B, k, c, p, c); k += 1; }
            else if (blocks[k] == 2) { int64_t p = pivots[k] - 1; if (p != k + 1) swap_el<N>(B, k + 1, c, p, c); k += 2; }
            else { *bad = 1; return; }
        }
    } else {
        int64_t k = n - 1;
        while (k >= 0) {
            if (blocks[k] == 1) { int64_t p = pivots[k] - 1; if (p != k) swap_el<N>(B, k, c, p, c); k -= 1; }
            else if (blocks[k] == 0 && k > 0 && blocks[k - 1] == 2) {
                int64_t s0 = k - 1, p = pivots[s0] - 1;
                if (p != s0 + 1) swap_el<N>(B, s0 + 1, c, p, c);
                k -= 2;
            } else { *bad = 1; return; }
        }
    }
}

// solve.jl:268-296: block-diagonal solve, one thread per (block start, column).
template <int N>
__global__ void k_ldlt_dsolve(DV B, int64_t n, int64_t nrhs, DV F, DV dsub, const uint8_t *blocks, Status *st) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * nrhs) return;
    const int64_t k = e % n, c = e / n;
    if (blocks[k] == 1) {
        stv<N>(B, k, c, mf::div(ld<N>(B, k, c), ld<N>(F, k, k)));
    } else if (blocks[k] == 2) {
        mf::V<N> s1, s2;
        if (!solve2x2<N>(ld<N>(F, k, k), ld<N>(dsub, k, 0), ld<N>(F, k + 1, k + 1), ld<N>(B, k, c), ld<N>(B, k + 1, c), s1, s2)) {
            st->info = k + 1;
            return;
        }
        stv<N>(B, k, c, s1);
        stv<N>(B, k + 1, c, s2);
    }
}

// Column-wise max |A| over rows [rlo(c), m): rlo = c for the lower triangle.
template <int N>
__global__ void k_col_absmax_region(DV A, int64_t m, int lower_only, DV out) {
    const int64_t c = blockIdx.x;
    __shared__ double sv_buf[256 * N];
    mf::V<N> mx = mf::zero<N>();
    // lower_only: 0 all rows, 1 rows >= c, 2 rows <= c
    const int64_t lo = lower_only == 1 ? c : 0;
    const int64_t hi = lower_only == 2 ? (c + 1 < m ? c + 1 : m) : m;
    for (int64_t r = lo + threadIdx.x; r < hi; r += blockDim.x) mx = mf::vmax(mx, mf::vabs(ld<N>(A, r, c)));
    mx = wg_max<N>(mx, sv_buf);
    if (threadIdx.x == 0) stv<N>(out, c, 0, mx);
}

// max over count column maxima (limb-planar count x 1) into out[0..N): the
// result of the sequential fold mx = vmax(mx, v_c), c ascending, from zero
// (the host reduction of the same values). One wave: each lane folds a
// contiguous chunk in order, then lane 0 folds the non-empty chunks' results
// in order; vmax keeps the later of two equal values, so this is the
// sequential fold's value, bit for bit.
template <int N>
__global__ __launch_bounds__(64) void k_absmax_fold(DV colmax, int64_t count, double *out) {
    __shared__ double part[64 * N];
    const int lane = threadIdx.x;
    const int64_t chunk = (count + 63) / 64;
    const int64_t lo = lane * chunk, hi = lo + chunk < count ? lo + chunk : count;
    mf::V<N> mx = mf::zero<N>();
    for (int64_t c = lo; c < hi; ++c) mx = mf::vmax(mx, ld<N>(colmax, c, 0));
    lds_put<N>(&part[lane * N], mx);
    __syncthreads();
    if (lane == 0) {
        mf::V<N> r = mf::zero<N>();
        for (int l = 0; l < 64 && (int64_t)l * chunk < count; ++l) r = mf::vmax(r, lds_get<N>(&part[l * N]));
#pragma unroll
        for (int q = 0; q < N; ++q) out[q] = r.x[q];
    }
}

__global__ void k_iota1(int64_t *p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i + 1;
}
__global__ void k_fill_u8(uint8_t *p, int64_t n, uint8_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ----------------------------------------------------------- launchers
#define MF_DISPATCH(N, ...)                                        \
    switch (N) {                                                   \
        case 1: { constexpr int NN = 1; __VA_ARGS__; } break;      \
        case 2: { constexpr int NN = 2; __VA_ARGS__; } break;      \
        case 3: { constexpr int NN = 3; __VA_ARGS__; } break;      \
        case 4: { constexpr int NN = 4; __VA_ARGS__; } break;      \
        default: return hipErrorInvalidValue;                      \
    }
static inline unsigned nb(int64_t work, int b) { return (unsigned)cdiv(work > 0 ? work : 1, b); }
template <int N> static S<N> mk(const double *v) { S<N> s; for (int l = 0; l < N; ++l) s.x[l] = v[l]; return s; }

hipError_t chol_panel(int N, hipStream_t s, const DV &A, int64_t first, int64_t last, Status *st) {
    if (last - first + 1 <= 16) {
        MF_DISPATCH(N, hipLaunchKernelGGL((k_chol_panel_lds<NN, 16>), dim3(1), dim3(256), 0, s, A, first, last, st));
    } else if (last - first + 1 <= CHOL_LDS) {
        MF_DISPATCH(N, hipLaunchKernelGGL((k_chol_panel_lds<NN, CHOL_LDS>), dim3(1), dim3(256), 0, s, A, first, last, st));
    } else {
        MF_DISPATCH(N, hipLaunchKernelGGL(k_chol_panel<NN>, dim3(1), dim3(256), 0, s, A, first, last, st));
    }
    return hipGetLastError();
}

hipError_t lu_pivot_swap(int N, hipStream_t s, const DV &A, int64_t k, int64_t m, int64_t c0, int64_t c1, int64_t *ipiv, Status *st) {
    if (m - k > INT32_MAX / 2) return hipErrorInvalidValue;
    MF_DISPATCH(N, hipLaunchKernelGGL(k_lu_pivot_swap<NN>, dim3(1), dim3(LP_THREADS), 0, s, A, k, m, c0, c1, ipiv, st));
    return hipGetLastError();
}

hipError_t lu_laswp(int N, hipStream_t s, const DV &A, int64_t first, int64_t count, const int64_t *ipiv, int64_t a0, int64_t acount,
                    int64_t b0, int64_t bcount) {
    if (acount + bcount <= 0 || count <= 0) return hipSuccess;
    MF_DISPATCH(N, hipLaunchKernelGGL(k_lu_laswp<NN>, dim3(nb((acount + bcount) * N, 256)), dim3(256), 0, s, A, first, count, ipiv, a0, acount,
                                      b0, bcount));
    return hipGetLastError();
}

hipError_t lu_scale_update(int N, hipStream_t s, const DV &A, int64_t k, int64_t last, int64_t m, Status *st) {
    const int64_t rows = m - k - 1;
    MF_DISPATCH(N, hipLaunchKernelGGL(k_lu_scale_update<NN>, dim3(nb((rows > 0 ? rows : 1) * LS_LANES, 256)), dim3(256), 0, s, A, k, last, m, st));
    return hipGetLastError();
}

hipError_t apply_ipiv(int N, hipStream_t s, const DV &B, int64_t count, int64_t nrhs, const int64_t *ipiv) {
    if (nrhs == 0 || count == 0) return hipSuccess;
    MF_DISPATCH(N, hipLaunchKernelGGL(k_apply_ipiv<NN>, dim3(nb(nrhs, 64)), dim3(64), 0, s, B, count, nrhs, ipiv));
    return hipGetLastError();
}

hipError_t ldlt_unblocked(int N, hipStream_t s, const DV &A, int64_t n, const DV &dsub, int64_t *pivots, uint8_t *blocks,
                          const double *alpha, Status *st) {
    MF_DISPATCH(N, hipLaunchKernelGGL(k_ldlt_unblocked<NN>, dim3(1), dim3(512), 0, s, A, n, dsub, pivots, blocks, mk<NN>(alpha), st));
    return hipGetLastError();
}

hipError_t ldlt_panel(int N, hipStream_t s, const DV &A, int64_t n, int64_t pfirst, int64_t reqlast, const DV &dsub, int64_t *pivots,
                      uint8_t *blocks, const double *alpha, double *keepbuf, int spec, Status *st) {
    MF_DISPATCH(N, hipLaunchKernelGGL(k_ldlt_panel<NN>, dim3(1), dim3(512), 0, s, A, n, pfirst, reqlast, dsub, pivots, blocks,
                                      mk<NN>(alpha), keepbuf, spec, st));
    return hipGetLastError();
}

size_t ldlt_panel_grid_doubles(int N, int64_t n) { return 4 * (size_t)N * (size_t)n + ldlt_grid_partials(N); }

hipError_t ldlt_panel_grid(int N, hipStream_t s, const DV &A, int64_t n, int64_t pfirst, int64_t reqlast, const DV &dsub, int64_t *pivots,
                           uint8_t *blocks, const double *alpha, double *buf, int spec, int trunc, int64_t weighted_ld, Status *st,
                           int workgroups) {
    if (workgroups < 1 || workgroups > GP_MAXG) return hipErrorInvalidValue;
    double *part = buf + 4 * (size_t)N * (size_t)n;
    auto launch = [&](auto ntag) -> hipError_t {
        constexpr int NN = decltype(ntag)::value;
        DV a = A, d = dsub;
        int64_t n_ = n, pf = pfirst, rl = reqlast;
        S<NN> al = mk<NN>(alpha);
        int sp = spec, tr = trunc;
        int64_t wl = weighted_ld;
        // an ordinary launch: the panel's workgroups (at most 64, one per CU)
        // are resident together on a device that runs nothing else meanwhile
        // (a cooperative launch costs ~11 us before and after the kernel);
        // should they not be, the bounded barrier ends the call with an error
        hipLaunchKernelGGL(k_ldlt_panel_grid<NN>, dim3((unsigned)workgroups), dim3(GP_THREADS), 0, s, a, n_, pf, rl, d, pivots, blocks, al, buf,
                           part, sp, tr, wl, st);
        return hipGetLastError();
    };
    switch (N) {
        case 1: return launch(std::integral_constant<int, 1>());
        case 2: return launch(std::integral_constant<int, 2>());
        case 3: return launch(std::integral_constant<int, 3>());
        case 4: return launch(std::integral_constant<int, 4>());
        default: return hipErrorInvalidValue;
    }
}

hipError_t ldlt_weighted(int N, hipStream_t s, const DV &A, int64_t pfirst, int64_t plast, int64_t tfirst, int64_t tcount,
                         const DV &dsub, const uint8_t *blocks, const DV &W, const Status *st) {
    if (tcount <= 0) return hipSuccess;
    MF_DISPATCH(N, hipLaunchKernelGGL(k_ldlt_weighted<NN>, dim3(nb(tcount, 256)), dim3(256), 0, s, A, pfirst, plast, tfirst, tcount,
                                      dsub, blocks, W, st));
    return hipGetLastError();
}

hipError_t mirror_lower(int N, hipStream_t s, const DV &A, int64_t n, const Status *st) {
    if (n <= 1) return hipSuccess;
    const unsigned t = (unsigned)((n + MIR_T - 1) / MIR_T);
    MF_DISPATCH(N, hipLaunchKernelGGL(k_mirror_lower<NN>, dim3(t, t), dim3(256), 0, s, A, n, st));
    return hipGetLastError();
}

hipError_t ldlt_restore(int N, hipStream_t s, const DV &A, int64_t n, int64_t pfirst, int64_t kstop, const double *keepbuf) {
    if (kstop <= pfirst) return hipSuccess;
    const unsigned bx = (unsigned)std::min<int64_t>(nb(n, 256), 64);
    MF_DISPATCH(N, hipLaunchKernelGGL(k_ldlt_restore<NN>, dim3(bx, (unsigned)(kstop - pfirst)), dim3(256), 0, s, A, n, pfirst, keepbuf));
    return hipGetLastError();
}

hipError_t ldlt_pivots(int N, hipStream_t s, const DV &B, int64_t n, int64_t nrhs, const int64_t *pivots, const uint8_t *blocks,
                       int forward, int *bad) {
    if (nrhs == 0 || n == 0) return hipSuccess;
    MF_DISPATCH(N, hipLaunchKernelGGL(k_ldlt_pivots<NN>, dim3(nb(nrhs, 64)), dim3(64), 0, s, B, n, nrhs, pivots, blocks, forward, bad));
    return hipGetLastError();
}

hipError_t ldlt_dsolve(int N, hipStream_t s, const DV &B, int64_t n, int64_t nrhs, const DV &F, const DV &dsub, const uint8_t *blocks,
                       Status *st) {
    if (nrhs == 0 || n == 0) return hipSuccess;
    MF_DISPATCH(N, hipLaunchKernelGGL(k_ldlt_dsolve<NN>, dim3(nb(n * nrhs, 256)), dim3(256), 0, s, B, n, nrhs, F, dsub, blocks, st));
    return hipGetLastError();
}

hipError_t col_absmax_region(int N, hipStream_t s, const DV &A, int64_t m, int64_t ncols, int lower_only, const DV &out) {
    if (ncols == 0) return hipSuccess;
    MF_DISPATCH(N, hipLaunchKernelGGL(k_col_absmax_region<NN>, dim3((unsigned)ncols), dim3(256), 0, s, A, m, lower_only, out));
    return hipGetLastError();
}

hipError_t absmax_fold(int N, hipStream_t s, const DV &colmax, int64_t count, double *out) {
    MF_DISPATCH(N, hipLaunchKernelGGL(k_absmax_fold<NN>, dim3(1), dim3(64), 0, s, colmax, count, out));
    return hipGetLastError();
}

hipError_t iota1(hipStream_t s, int64_t *p, int64_t n) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_iota1, dim3(nb(n, 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}
hipError_t fill_u8(hipStream_t s, uint8_t *p, int64_t n, uint8_t v) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u8, dim3(nb(n, 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

}  // namespace mfk
