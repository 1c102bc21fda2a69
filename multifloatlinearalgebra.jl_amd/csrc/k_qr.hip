// Householder QR kernels: qr!, rrqr! (both routes) and apply_q!
// (src/factorizations/qr.jl:87-956), in MFLA's per-element order.
//
// Every reduction in MFLA's QR is a serial chain in ascending row order (a
// column norm, a reflector projection, a gemv entry). One wave owns one such
// chain: its 64 lanes evaluate 64 consecutive terms at once (products,
// absolute values, the divisions of the scaled norm), and the chain itself
// runs on the wave with each lane's term broadcast by readlane, so only the
// additions are serial and every lane ends with the (uniform) result.
//
// The scaled column norm (qr.jl:96-122) keeps scale = the largest |a| seen
// so far; lane r's ratio needs only the maximum over the rows before it, an
// exclusive prefix maximum that the wave forms with shuffles (ties keep the
// earlier row, as the serial scan does). The divisions then run in parallel
// and the serial part is one Float64xN add (or two products and an add when
// the scale grows).
#include "mf_common.hpp"
#include "k_qr.hpp"

namespace mfk {

namespace {

__device__ __forceinline__ double lane_bcast(double v, int k) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, k);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), k);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <int N> __device__ __forceinline__ mf::V<N> bcast(const mf::V<N> &v, int k) {
    mf::V<N> r;
#pragma unroll
    for (int l = 0; l < N; ++l) r.x[l] = lane_bcast(v.x[l], k);
    return r;
}

template <int N> __device__ __forceinline__ mf::V<N> shfl_up_v(const mf::V<N> &v, int off) {
    mf::V<N> r;
#pragma unroll
    for (int l = 0; l < N; ++l) r.x[l] = __shfl_up(v.x[l], (unsigned)off, 64);
    return r;
}

// acc + sum_{r < count} prod(r), added in ascending r (the whole wave calls it
// with uniform arguments; the result is uniform).
template <int N, class Prod> __device__ mf::V<N> wave_dot(mf::V<N> acc, int64_t count, Prod prod) {
    const int lane = threadIdx.x & 63;
    for (int64_t base = 0; base < count; base += 64) {
        const int64_t r = base + lane;
        const mf::V<N> p = r < count ? prod(r) : mf::zero<N>();
        const int lim = (int)((count - base) < 64 ? (count - base) : 64);
        for (int k = 0; k < lim; ++k) acc = mf::add(acc, bcast<N>(p, k));
    }
    return acc;
}

// The scaled norm state (scale, scaled_sum) of value(0 .. count-1), qr.jl:96-122.
template <int N, class Value> __device__ void wave_norm(int64_t count, Value value, mf::V<N> &scale, mf::V<N> &ssum) {
    const int lane = threadIdx.x & 63;
    const mf::V<N> one = mf::from<N>(1.0), zero = mf::zero<N>();
    mf::V<N> run = zero;
    ssum = one;
    for (int64_t base = 0; base < count; base += 64) {
        const int64_t r = base + lane;
        const mf::V<N> a = r < count ? mf::vabs(value(r)) : zero;
        mf::V<N> inc = a;                                   // max over lanes 0..lane, earlier wins ties
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const mf::V<N> o = shfl_up_v<N>(inc, off);
            if (lane >= off && !mf::gt(inc, o)) inc = o;
        }
        mf::V<N> ex = shfl_up_v<N>(inc, 1);                 // max over the rows before r
        if (lane == 0 || !mf::gt(ex, run)) ex = run;
        int kind = 0;
        mf::V<N> q = zero;
        if (r < count && !mf::is_zero(a)) {
            if (mf::is_zero(ex)) kind = 1;
            else if (mf::gt(a, ex)) { kind = 2; q = mf::div(ex, a); }
            else { kind = 3; const mf::V<N> t = mf::div(a, ex); q = mf::mul(t, t); }
        }
        const int lim = (int)((count - base) < 64 ? (count - base) : 64);
        for (int k = 0; k < lim; ++k) {
            const int kk = __builtin_amdgcn_readlane(kind, k);
            if (kk == 0) continue;
            const mf::V<N> qk = bcast<N>(q, k);
            if (kk == 1) ssum = one;
            else if (kk == 2) ssum = mf::add(one, mf::mul(mf::mul(ssum, qk), qk));
            else ssum = mf::add(ssum, qk);
        }
        const mf::V<N> last = bcast<N>(inc, lim - 1);
        if (mf::gt(last, run)) run = last;
    }
    scale = run;
}

template <int N> __device__ __forceinline__ mf::V<N> norm_of(const mf::V<N> &scale, const mf::V<N> &ssum) {
    return mf::is_zero(scale) ? mf::zero<N>() : mf::mul(scale, mf::sqrt(ssum));
}

template <int N> __device__ __forceinline__ mf::V<N> vec_ld(const double *p, int64_t i, int64_t ls) {
    mf::V<N> r;
#pragma unroll
    for (int l = 0; l < N; ++l) r.x[l] = p[l * ls + i];
    return r;
}
template <int N> __device__ __forceinline__ void vec_st(double *p, int64_t i, int64_t ls, const mf::V<N> &v) {
#pragma unroll
    for (int l = 0; l < N; ++l) p[l * ls + i] = v.x[l];
}

template <int N> __device__ __forceinline__ bool le(const mf::V<N> &x, const mf::V<N> &y) { return mf::lt(x, y) || mf::eq(x, y); }

// F(p, c): QR_PANEL x n, limb stride QR_PANEL * n
template <int N> __device__ __forceinline__ mf::V<N> f_ld(const double *F, int64_t n, int64_t p, int64_t c) {
    return vec_ld<N>(F, c * QR_PANEL + p, QR_PANEL * n);
}
template <int N> __device__ __forceinline__ void f_st(double *F, int64_t n, int64_t p, int64_t c, const mf::V<N> &v) {
    vec_st<N>(F, c * QR_PANEL + p, QR_PANEL * n, v);
}

constexpr int WAVES = 4;   // waves per 256-thread workgroup

}  // namespace

// ------------------------------------------------------------ norm state
template <int N>
__global__ __launch_bounds__(256) void k_qr_norms(int64_t m, int64_t n, DV A, int64_t first_row, int64_t c0, int64_t ncols, QrState q) {
    const int64_t w = (int64_t)blockIdx.x * WAVES + (threadIdx.x >> 6);
    if (w >= ncols) return;                                   // whole wave
    const int64_t c = c0 + w;
    mf::V<N> sc, ss;
    wave_norm<N>(m - first_row, [&](int64_t r) { return ld<N>(A, first_row + r, c); }, sc, ss);
    if ((threadIdx.x & 63) == 0) {
        vec_st<N>(q.scale, c, n, sc);
        vec_st<N>(q.sum, c, n, ss);
        q.dirty[c] = 0;
    }
}

// ------------------------------------------------------------- reflector
template <int N>
__global__ __launch_bounds__(256) void k_qr_reflector(int64_t m, int64_t step, DV A, DV tau, int blocked, QrState q) {
    __shared__ double lead[N];
    __shared__ int nonzero;
    if (threadIdx.x < 64) {
        mf::V<N> sc, ss;
        wave_norm<N>(m - step, [&](int64_t r) { return ld<N>(A, step + r, step); }, sc, ss);
        const mf::V<N> nrm = norm_of<N>(sc, ss);
        if (threadIdx.x == 0) {
            const mf::V<N> alpha = ld<N>(A, step, step);
            mf::V<N> t = mf::zero<N>(), diag = alpha;
            nonzero = !mf::is_zero(nrm);
            if (nonzero) {
                const mf::V<N> beta = mf::ge(alpha, mf::zero<N>()) ? mf::neg(nrm) : nrm;
                const mf::V<N> leading = mf::sub(alpha, beta);
                t = mf::div(mf::sub(beta, alpha), beta);
#pragma unroll
                for (int l = 0; l < N; ++l) lead[l] = leading.x[l];
                diag = beta;
            }
            stv<N>(tau, step, 0, t);
            if (blocked) {
                vec_st<N>(q.diag, 0, 1, diag);
                stv<N>(A, step, step, mf::from<N>(1.0));
            } else {
                stv<N>(A, step, step, diag);
            }
        }
    }
    __syncthreads();
    if (!nonzero) return;
    mf::V<N> leading;
#pragma unroll
    for (int l = 0; l < N; ++l) leading.x[l] = lead[l];
    for (int64_t r = step + 1 + threadIdx.x; r < m; r += blockDim.x) stv<N>(A, r, step, mf::div(ld<N>(A, r, step), leading));
}

// -------------------------------------------------------- apply reflector
template <int N>
__global__ __launch_bounds__(256) void k_qr_apply(int64_t m, int64_t step, DV V, int64_t vcol, DV T, int64_t c0, int64_t ncols, DV tau) {
    const int64_t w = (int64_t)blockIdx.x * WAVES + (threadIdx.x >> 6);
    if (w >= ncols) return;
    const mf::V<N> t = ld<N>(tau, step, 0);
    if (mf::is_zero(t)) return;
    const int64_t c = c0 + w;
    const int lane = threadIdx.x & 63;
    mf::V<N> proj = wave_dot<N>(ld<N>(T, step, c), m - step - 1,
                                [&](int64_t r) { return mf::mul(ld<N>(V, step + 1 + r, vcol), ld<N>(T, step + 1 + r, c)); });
    proj = mf::mul(proj, t);
    if (lane == 0) stv<N>(T, step, c, mf::sub(ld<N>(T, step, c), proj));
    for (int64_t r = step + 1 + lane; r < m; r += 64) stv<N>(T, r, c, mf::sub(ld<N>(T, r, c), mf::mul(ld<N>(V, r, vcol), proj)));
}

// ------------------------------------------------------------ pivot search
// qr.jl:192-224 with the rebuilds done ahead in parallel: pivot_norm only
// grows during the scan, so every column the serial scan rebuilds satisfies
// the rebuild test against the first pivot norm; those candidates are rebuilt
// by all waves into tscale / tsum, and the serial scan (wave 0) commits a
// rebuild exactly where MFLA's scan performs it (rebuilding inline if a
// candidate was not prepared).
template <int N>
__global__ __launch_bounds__(1024) void k_qr_pivot_hybrid(int64_t m, int64_t n, int64_t step, DV A, const int64_t *perm, QrState q) {
    __shared__ double pn0s[N];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    const mf::V<N> onem = mf::sub(mf::from<N>(1.0), mf::from<N>(ldexp(1.0, -26 * N)));
    if (wave == 0) {
        mf::V<N> sc, ss;
        wave_norm<N>(m - step, [&](int64_t r) { return ld<N>(A, step + r, step); }, sc, ss);
        const mf::V<N> pn = norm_of<N>(sc, ss);
        if (lane == 0) {
            vec_st<N>(q.scale, step, n, sc);
            vec_st<N>(q.sum, step, n, ss);
            q.dirty[step] = 0;
#pragma unroll
            for (int l = 0; l < N; ++l) pn0s[l] = pn.x[l];
        }
    }
    __syncthreads();
    mf::V<N> pn0;
#pragma unroll
    for (int l = 0; l < N; ++l) pn0.x[l] = pn0s[l];
    const mf::V<N> thr0 = mf::mul(onem, pn0);
    for (int64_t c = step + 1 + wave; c < n; c += nw) {
        const mf::V<N> cand = norm_of<N>(vec_ld<N>(q.scale, c, n), vec_ld<N>(q.sum, c, n));
        const bool need = q.dirty[c] || mf::is_zero(pn0) || mf::ge(cand, thr0);
        if (need) {
            mf::V<N> sc, ss;
            wave_norm<N>(m - step, [&](int64_t r) { return ld<N>(A, step + r, c); }, sc, ss);
            if (lane == 0) { vec_st<N>(q.tscale, c, n, sc); vec_st<N>(q.tsum, c, n, ss); }
        }
        if (lane == 0) q.tflag[c] = need ? 1 : 0;
    }
    __syncthreads();
    if (wave != 0) return;
    int64_t pivot = step;
    mf::V<N> pn = pn0;
    for (int64_t c = step + 1; c < n; ++c) {
        mf::V<N> sc = vec_ld<N>(q.scale, c, n), ss = vec_ld<N>(q.sum, c, n);
        mf::V<N> cand = norm_of<N>(sc, ss);
        if (q.dirty[c] || mf::is_zero(pn) || mf::ge(cand, mf::mul(onem, pn))) {
            if (q.tflag[c]) {
                sc = vec_ld<N>(q.tscale, c, n);
                ss = vec_ld<N>(q.tsum, c, n);
            } else {
                wave_norm<N>(m - step, [&](int64_t r) { return ld<N>(A, step + r, c); }, sc, ss);
            }
            if (lane == 0) { vec_st<N>(q.scale, c, n, sc); vec_st<N>(q.sum, c, n, ss); q.dirty[c] = 0; }
            cand = norm_of<N>(sc, ss);
        }
        if (mf::gt(cand, pn) || (mf::eq(cand, pn) && perm[c] < perm[pivot])) { pivot = c; pn = cand; }
    }
    if (lane == 0) *q.pivot = pivot;
}

// qr.jl:344-362: argmax of the downdated norms, ties to the smaller original
// column (perm); 256 threads, thread-local scans then a tree.
template <int N>
__global__ __launch_bounds__(256) void k_qr_pivot_downdated(int64_t n, int64_t step, const int64_t *perm, QrState q) {
    __shared__ double bn[256 * N];
    __shared__ int64_t bc[256];
    const int t = threadIdx.x;
    auto better = [&](const mf::V<N> &a, int64_t ca, const mf::V<N> &b, int64_t cb) {   // is a before b?
        return mf::gt(a, b) || (mf::eq(a, b) && perm[ca] < perm[cb]);
    };
    int64_t best = -1;
    mf::V<N> bv = mf::zero<N>();
    for (int64_t c = step + t; c < n; c += blockDim.x) {
        const mf::V<N> v = norm_of<N>(vec_ld<N>(q.scale, c, n), vec_ld<N>(q.sum, c, n));
        if (best < 0 || better(v, c, bv, best)) { best = c; bv = v; }
    }
#pragma unroll
    for (int l = 0; l < N; ++l) bn[l * 256 + t] = bv.x[l];
    bc[t] = best;
    __syncthreads();
    for (int h = 128; h >= 1; h >>= 1) {
        if (t < h) {
            const int64_t co = bc[t + h];
            if (co >= 0) {
                mf::V<N> vo;
#pragma unroll
                for (int l = 0; l < N; ++l) vo.x[l] = bn[l * 256 + t + h];
                mf::V<N> vm;
#pragma unroll
                for (int l = 0; l < N; ++l) vm.x[l] = bn[l * 256 + t];
                if (bc[t] < 0 || better(vo, co, vm, bc[t])) {
#pragma unroll
                    for (int l = 0; l < N; ++l) bn[l * 256 + t] = vo.x[l];
                    bc[t] = co;
                }
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        *q.pivot = bc[0];
        *q.unreliable = 0;
    }
}

// ------------------------------------------------------------------ swap
template <int N>
__global__ void k_qr_swap(int64_t m, int64_t n, int64_t step, int64_t ls, DV A, int64_t *perm, QrState q) {
    const int64_t p = *q.pivot;
    if (p == step) return;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < m) {
        const mf::V<N> a = ld<N>(A, r, step), b = ld<N>(A, r, p);
        stv<N>(A, r, step, b);
        stv<N>(A, r, p, a);
    }
    if (r < ls) {
        const mf::V<N> a = f_ld<N>(q.F, n, r, step), b = f_ld<N>(q.F, n, r, p);
        f_st<N>(q.F, n, r, step, b);
        f_st<N>(q.F, n, r, p, a);
    }
    if (r == 0) {
        const int64_t t = perm[step]; perm[step] = perm[p]; perm[p] = t;
        const mf::V<N> s0 = vec_ld<N>(q.scale, step, n), s1 = vec_ld<N>(q.scale, p, n);
        vec_st<N>(q.scale, step, n, s1); vec_st<N>(q.scale, p, n, s0);
        const mf::V<N> u0 = vec_ld<N>(q.sum, step, n), u1 = vec_ld<N>(q.sum, p, n);
        vec_st<N>(q.sum, step, n, u1); vec_st<N>(q.sum, p, n, u0);
        const int d = q.dirty[step]; q.dirty[step] = q.dirty[p]; q.dirty[p] = d;
    }
}

// ------------------------------------------------------ blocked route steps
// qr.jl:527-538: A(r, step) -= sum_prior A(r, bs+prior) F(prior, step)
template <int N>
__global__ void k_qr_delayed_column(int64_t m, int64_t n, int64_t step, int64_t bs, int64_t ls, DV A, QrState q) {
    const int64_t r = step + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    mf::V<N> corr = mf::zero<N>();
    for (int64_t p = 0; p < ls; ++p) corr = mf::add(corr, mf::mul(ld<N>(A, r, bs + p), f_ld<N>(q.F, n, p, step)));
    stv<N>(A, r, step, mf::sub(ld<N>(A, r, step), corr));
}

// qr.jl:556-565 gemv!: aux(p) = (-tau) * sum_r A(r, bs+p) A(r, step), r >= step
template <int N>
__global__ __launch_bounds__(256) void k_qr_aux(int64_t m, int64_t step, int64_t bs, int64_t ls, DV A, DV tau, QrState q) {
    const int64_t p = (int64_t)blockIdx.x * WAVES + (threadIdx.x >> 6);
    if (p >= ls) return;
    const mf::V<N> acc = wave_dot<N>(mf::zero<N>(), m - step, [&](int64_t r) { return mf::mul(ld<N>(A, step + r, bs + p), ld<N>(A, step + r, step)); });
    if ((threadIdx.x & 63) == 0) vec_st<N>(q.aux, p, QR_PANEL, mf::mul(mf::neg(ld<N>(tau, step, 0)), acc));
}

// qr.jl:545-585 for column c > step: the gemv! entry F(ls, c) = tau * sum_r
// A(r, c) A(r, step), its delayed correction sum_prior F(prior, c) aux(prior),
// then the materialised R entry A(step, c) -= sum_{p <= ls} A(step, bs+p) F(p, c).
template <int N>
__global__ __launch_bounds__(256) void k_qr_frow(int64_t m, int64_t n, int64_t step, int64_t bs, int64_t ls, DV A, DV tau, QrState q) {
    const int64_t c = step + 1 + (int64_t)blockIdx.x * WAVES + (threadIdx.x >> 6);
    if (c >= n) return;
    const mf::V<N> acc = wave_dot<N>(mf::zero<N>(), m - step, [&](int64_t r) { return mf::mul(ld<N>(A, step + r, c), ld<N>(A, step + r, step)); });
    mf::V<N> f = mf::mul(ld<N>(tau, step, 0), acc);
    if (ls > 0) {
        mf::V<N> corr = mf::zero<N>();
        for (int64_t p = 0; p < ls; ++p) corr = mf::add(corr, mf::mul(f_ld<N>(q.F, n, p, c), vec_ld<N>(q.aux, p, QR_PANEL)));
        f = mf::add(f, corr);
    }
    mf::V<N> corr = mf::zero<N>();
    for (int64_t p = 0; p <= ls; ++p) corr = mf::add(corr, mf::mul(ld<N>(A, step, bs + p), p == ls ? f : f_ld<N>(q.F, n, p, c)));
    if ((threadIdx.x & 63) == 0) {
        f_st<N>(q.F, n, ls, c, f);
        stv<N>(A, step, c, mf::sub(ld<N>(A, step, c), corr));
    }
}

// ------------------------------------------------------------ norm update
template <int N>
__global__ __launch_bounds__(256) void k_qr_norm_update(int64_t m, int64_t n, int64_t step, DV A, DV tau, int64_t bs, int64_t pc,
                                                        int restore_diag, QrState q) {
    if (restore_diag && blockIdx.x == 0 && threadIdx.x == 0) stv<N>(A, step, step, vec_ld<N>(q.diag, 0, 1));
    const int64_t c = step + 1 + (int64_t)blockIdx.x * WAVES + (threadIdx.x >> 6);
    if (c >= n) return;
    const int lane = threadIdx.x & 63;
    const mf::V<N> floor_ = mf::from<N>(ldexp(1.0, 4 - 52 * N));
    if (!mf::is_zero(ld<N>(tau, step, 0)) && !q.dirty[c]) {
        const mf::V<N> cs = vec_ld<N>(q.scale, c, n), removed = mf::vabs(ld<N>(A, step, c));
        if (mf::is_zero(cs)) {
            if (lane == 0) q.dirty[c] = mf::is_zero(removed) ? 0 : 1;
        } else {
            const mf::V<N> ratio = mf::div(removed, cs);
            const mf::V<N> ns = mf::sub(vec_ld<N>(q.sum, c, n), mf::mul(ratio, ratio));
            if (!mf::is_finite(ns) || le<N>(ns, floor_)) {
                mf::V<N> sc, ss;
                wave_norm<N>(m - step - 1, [&](int64_t r) {
                    mf::V<N> v = ld<N>(A, step + 1 + r, c);
                    for (int64_t p = 0; p < pc; ++p) v = mf::sub(v, mf::mul(ld<N>(A, step + 1 + r, bs + p), f_ld<N>(q.F, n, p, c)));
                    return v;
                }, sc, ss);
                if (lane == 0) { vec_st<N>(q.scale, c, n, sc); vec_st<N>(q.sum, c, n, ss); q.dirty[c] = 0; }
            } else if (lane == 0) {
                vec_st<N>(q.sum, c, n, ns);
            }
        }
    }
    if (lane == 0 && q.dirty[c]) atomicOr(q.unreliable, 1);
}

template <int N>
__global__ void k_qr_panel_reset(int64_t n, QrState q) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < (int64_t)N * QR_PANEL * n) q.F[e] = 0.0;
    if (e == 0) *q.unreliable = 0;
}

// ---------------------------------------------------------------- launchers
#define QR_DISPATCH(N, ...)                                      \
    switch (N) {                                                 \
        case 1: { constexpr int NN = 1; __VA_ARGS__; } break;    \
        case 2: { constexpr int NN = 2; __VA_ARGS__; } break;    \
        case 3: { constexpr int NN = 3; __VA_ARGS__; } break;    \
        case 4: { constexpr int NN = 4; __VA_ARGS__; } break;    \
        default: return hipErrorInvalidValue;                    \
    }                                                            \
    return hipGetLastError();

static dim3 wave_grid(int64_t waves) { return dim3((unsigned)cdiv(waves, WAVES)); }

hipError_t qr_norms(int N, hipStream_t s, int64_t m, int64_t n, const DV &A, int64_t first_row, int64_t c0, int64_t ncols, const QrState &q) {
    if (ncols <= 0) return hipSuccess;
    QR_DISPATCH(N, hipLaunchKernelGGL(k_qr_norms<NN>, wave_grid(ncols), dim3(256), 0, s, m, n, A, first_row, c0, ncols, q));
}

hipError_t qr_reflector(int N, hipStream_t s, int64_t m, int64_t step, const DV &A, const DV &tau, int blocked, const QrState &q) {
    QR_DISPATCH(N, hipLaunchKernelGGL(k_qr_reflector<NN>, dim3(1), dim3(256), 0, s, m, step, A, tau, blocked, q));
}

hipError_t qr_apply(int N, hipStream_t s, int64_t m, int64_t step, const DV &V, int64_t vcol, const DV &T, int64_t c0, int64_t ncols,
                    const DV &tau) {
    if (ncols <= 0) return hipSuccess;
    QR_DISPATCH(N, hipLaunchKernelGGL(k_qr_apply<NN>, wave_grid(ncols), dim3(256), 0, s, m, step, V, vcol, T, c0, ncols, tau));
}

hipError_t qr_pivot_hybrid(int N, hipStream_t s, int64_t m, int64_t n, int64_t step, const DV &A, const int64_t *perm, const QrState &q) {
    QR_DISPATCH(N, hipLaunchKernelGGL(k_qr_pivot_hybrid<NN>, dim3(1), dim3(1024), 0, s, m, n, step, A, perm, q));
}

hipError_t qr_pivot_downdated(int N, hipStream_t s, int64_t n, int64_t step, const int64_t *perm, const QrState &q) {
    QR_DISPATCH(N, hipLaunchKernelGGL(k_qr_pivot_downdated<NN>, dim3(1), dim3(256), 0, s, n, step, perm, q));
}

hipError_t qr_swap(int N, hipStream_t s, int64_t m, int64_t n, int64_t step, int64_t ls, const DV &A, int64_t *perm, const QrState &q) {
    const int64_t rows = m > ls ? m : ls;
    QR_DISPATCH(N, hipLaunchKernelGGL(k_qr_swap<NN>, dim3((unsigned)cdiv(rows, 256)), dim3(256), 0, s, m, n, step, ls, A, perm, q));
}

hipError_t qr_delayed_column(int N, hipStream_t s, int64_t m, int64_t n, int64_t step, int64_t bs, int64_t ls, const DV &A, const QrState &q) {
    if (ls <= 0 || step >= m) return hipSuccess;
    QR_DISPATCH(N, hipLaunchKernelGGL(k_qr_delayed_column<NN>, dim3((unsigned)cdiv(m - step, 256)), dim3(256), 0, s, m, n, step, bs, ls, A, q));
}

hipError_t qr_aux(int N, hipStream_t s, int64_t m, int64_t n, int64_t step, int64_t bs, int64_t ls, const DV &A, const DV &tau, const QrState &q) {
    (void)n;
    if (ls <= 0) return hipSuccess;
    QR_DISPATCH(N, hipLaunchKernelGGL(k_qr_aux<NN>, wave_grid(ls), dim3(256), 0, s, m, step, bs, ls, A, tau, q));
}

hipError_t qr_frow(int N, hipStream_t s, int64_t m, int64_t n, int64_t step, int64_t bs, int64_t ls, const DV &A, const DV &tau, const QrState &q) {
    if (step + 1 >= n) return hipSuccess;
    QR_DISPATCH(N, hipLaunchKernelGGL(k_qr_frow<NN>, wave_grid(n - step - 1), dim3(256), 0, s, m, n, step, bs, ls, A, tau, q));
}

hipError_t qr_norm_update(int N, hipStream_t s, int64_t m, int64_t n, int64_t step, const DV &A, const DV &tau, int64_t bs, int64_t pc,
                          int restore_diag, const QrState &q) {
    const int64_t cols = n - step - 1 > 0 ? n - step - 1 : 0;
    const dim3 g = wave_grid(cols > 0 ? cols : 1);
    QR_DISPATCH(N, hipLaunchKernelGGL(k_qr_norm_update<NN>, g, dim3(256), 0, s, m, n, step, A, tau, bs, pc, restore_diag, q));
}

hipError_t qr_panel_reset(int N, hipStream_t s, int64_t n, const QrState &q) {
    QR_DISPATCH(N, hipLaunchKernelGGL(k_qr_panel_reset<NN>, dim3((unsigned)cdiv((int64_t)N * QR_PANEL * n, 256)), dim3(256), 0, s, n, q));
}

}  // namespace mfk
