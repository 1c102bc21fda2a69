// Batched factor + solve of many small uneven systems in one persistent launch.
//
// MFLA has no batched API: callers loop cholesky!/ldlt! + ldiv! over systems.
// Here one workgroup owns one system at a time and pulls the next system from
// a device work queue (one atomic per system), so thousands of uneven systems
// run in a single launch that a HIP graph can capture (no host
// synchronisation, no allocation). Each system runs MFLA's own algorithm:
//   Cholesky: blocked right-looking, block 16 (cholesky.jl:68-107): panel
//             columns in order, right-side TRSM per row, SYRK trailing update
//             with alpha = -1, beta = 1 -- bitwise the single-matrix result;
//   LDLT:     unblocked Bunch-Kaufman (ldlt.jl:352-412, the route MFLA takes
//             for n < 512) -- bitwise, pivots / blocks / inertia exact.
// Solves: forward substitutions are systolic in MFLA's order (bitwise); the
// backward substitution is either MFLA's exact order (one lane per RHS) or
// the systolic order (one pivot per step, rounding within c*n*u).
#pragma once
#include "mf_common.hpp"
#include "k_factor_dev.hpp"
#include "k_batched.hpp"

namespace mfk {

constexpr int BT_MAX = 256;  // largest workgroup; systems need n <= its thread count
constexpr int RHS_MAX = 8;   // right-hand sides per system

template <int N>
__device__ __forceinline__ bool wg_all_finite_lower(const DV &A, int64_t n, int *flag) {
    if (threadIdx.x == 0) *flag = 1;
    __syncthreads();
    for (int64_t e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int64_t r = e % n, c = e / n;
        if (r >= c && !mf::is_finite(ld<N>(A, r, c))) *flag = 0;
    }
    __syncthreads();
    const bool ok = *flag != 0;
    __syncthreads();
    return ok;
}

// cholesky.jl:10-107 for one system, one workgroup.
template <int N>
__device__ __forceinline__ int64_t wg_potrf(const DV &A, int64_t n, int64_t block, double *piv_s, int *flag) {
    if (!wg_all_finite_lower<N>(A, n, flag)) return -1;
    for (int64_t first = 0; first < n; first += block) {
        const int64_t last = (first + block < n ? first + block : n) - 1;
        for (int64_t c = first; c <= last; ++c) {
            if (threadIdx.x == 0) {
                mf::V<N> d = ld<N>(A, c, c);
                for (int64_t k = first; k < c; ++k) { mf::V<N> v = ld<N>(A, c, k); d = mf::msub(d, v, v); }
                if (!mf::gt(d, mf::zero<N>())) {
                    *flag = (int)(c + 1);
                } else {
                    *flag = 0;
                    mf::V<N> p = mf::sqrt(d);
                    stv<N>(A, c, c, p);
                    lds_put<N>(piv_s, p);
                }
            }
            __syncthreads();
            const int fail = *flag;
            if (fail) return fail;
            const mf::V<N> p = lds_get<N>(piv_s);
            for (int64_t r = c + 1 + threadIdx.x; r <= last; r += blockDim.x) {
                mf::V<N> v = ld<N>(A, r, c);
                for (int64_t k = first; k < c; ++k) v = mf::msub(v, ld<N>(A, r, k), ld<N>(A, c, k));
                stv<N>(A, r, c, mf::div(v, p));
            }
            __syncthreads();
        }
        if (last + 1 < n) {
            // trsm!(A21, L11; side=:right, uplo=:lower, trans=:T): forward over panel columns per row
            for (int64_t r = last + 1 + threadIdx.x; r < n; r += blockDim.x) {
                for (int64_t p = first; p <= last; ++p) {
                    mf::V<N> v = ld<N>(A, r, p);
                    for (int64_t q = first; q < p; ++q) v = mf::msub(v, ld<N>(A, r, q), ld<N>(A, p, q));
                    stv<N>(A, r, p, mf::div(v, ld<N>(A, p, p)));
                }
            }
            __syncthreads();
            // syrk!(A22, A21', -1, 1) over the lower triangle
            const int64_t w = n - last - 1;
            const mf::V<N> mone = mf::from<N>(-1.0), one = mf::from<N>(1.0);
            for (int64_t e = threadIdx.x; e < w * w; e += blockDim.x) {
                const int64_t r = last + 1 + e % w, c = last + 1 + e / w;
                if (r < c) continue;
                mf::V<N> acc = mf::zero<N>();
                for (int64_t k = first; k <= last; ++k) acc = mf::mac(acc, ld<N>(A, r, k), ld<N>(A, c, k));
                stv<N>(A, r, c, mf::add(mf::mul(mone, acc), mf::mul(one, ld<N>(A, r, c))));
            }
            __syncthreads();
        }
    }
    return 0;
}

// Systolic triangular solve on up to RHS_MAX right-hand sides: thread p owns
// row p. Forward: rows subtract pivots in ascending order (MFLA order).
// Backward: rows subtract pivots in descending order. The right-hand sides go
// through in chunks of RHS_CHUNK, each chunk's values in registers.
constexpr int RHS_CHUNK = 4;

template <int N>
__device__ __forceinline__ void wg_systolic(const DV &X, const DV &CA, const DV &D, int64_t n, int nrhs, bool forward, bool unit, double *xs) {
    const int64_t p = threadIdx.x;
    for (int c0 = 0; c0 < nrhs; c0 += RHS_CHUNK) {
        const int nc = nrhs - c0 < RHS_CHUNK ? nrhs - c0 : RHS_CHUNK;
        mf::V<N> v[RHS_CHUNK];   // every index below is a compile-time constant
#pragma unroll
        for (int c = 0; c < RHS_CHUNK; ++c) v[c] = (p < n && c < nc) ? ld<N>(X, p, c0 + c) : mf::zero<N>();
        for (int64_t s = 0; s < n; ++s) {
            const int64_t q = forward ? s : n - 1 - s;
            if (p == q) {
#pragma unroll
                for (int c = 0; c < RHS_CHUNK; ++c) {
                    if (c < nc) {
                        if (!unit) v[c] = mf::div(v[c], ld<N>(D, q, q));
                        stv<N>(X, q, c0 + c, v[c]);
                        lds_put<N>(&xs[c * N], v[c]);
                    }
                }
            }
            __syncthreads();
            if (p < n && (forward ? p > q : p < q)) {
                const mf::V<N> a = ld<N>(CA, p, q);
#pragma unroll
                for (int c = 0; c < RHS_CHUNK; ++c)
                    if (c < nc) v[c] = mf::msub(v[c], a, lds_get<N>(&xs[c * N]));
            }
            __syncthreads();
        }
    }
}

// MFLA's exact backward order: one lane per right-hand side (trsm.jl:56-77).
template <int N>
__device__ __forceinline__ void wg_exact_backward(const DV &X, const DV &CA, const DV &D, int64_t n, int nrhs, bool unit) {
    const int c = threadIdx.x;
    if (c < nrhs) {
        for (int64_t p = n - 1; p >= 0; --p) {
            mf::V<N> v = ld<N>(X, p, c);
            for (int64_t q = p + 1; q < n; ++q) v = mf::msub(v, ld<N>(CA, p, q), ld<N>(X, q, c));
            if (!unit) v = mf::div(v, ld<N>(D, p, p));
            stv<N>(X, p, c, v);
        }
    }
    __syncthreads();
}

template <int N>
__device__ __forceinline__ void wg_backward(const DV &X, const DV &CA, const DV &D, int64_t n, int nrhs, bool unit, int exact, double *xs) {
    if (exact) wg_exact_backward<N>(X, CA, D, n, nrhs, unit);
    else wg_systolic<N>(X, CA, D, n, nrhs, false, unit, xs);
}

// The workgroup shape BT (64, 128 or 256 threads: one, two or four waves)
// bounds the order of the systems it takes (n <= BT: a thread per row) and
// sets how many systems share a CU: the same 16 waves per CU are 4, 8 or 16
// systems. Up to three limbs the kernel is held to 128 VGPRs: four waves per
// SIMD hide each other's per-column barriers and latencies better than two
// unspilled ones (10,000 x3 systems: 122 -> 100 ms, with a small spill).
template <int N, int BT>
__global__ __launch_bounds__(BT, N <= 3 ? 4 : 2) void k_batched(const BatchEntry *E, int64_t count, double *Aall, double *Ball, double *Dall,
                                               int64_t *Pall, uint8_t *Kall, int64_t *info, double *origmax, unsigned int *counter,
                                               S<N> alpha_s, int64_t chol_block, int exact_backward) {
    __shared__ double sv_buf[BT * N];
    __shared__ int64_t si_buf[BT];
    __shared__ double xs[RHS_MAX * N];
    __shared__ double piv_s[N];
    __shared__ double cbuf[2 * BT * N];
    __shared__ int flag;
    __shared__ int64_t next_s;
    __shared__ int64_t info_s;
    const mf::V<N> alpha = sv(alpha_s);
    while (true) {
        if (threadIdx.x == 0) next_s = (int64_t)atomicAdd(counter, 1u);
        __syncthreads();
        const int64_t s = next_s;
        __syncthreads();
        if (s >= count) break;
        const BatchEntry e = E[s];
        const int64_t n = e.n;
        const int nrhs = e.nrhs;
        if (n < 0 || n > BT || nrhs < 0 || nrhs > RHS_MAX) {   // unsupported shape: reported, never touched
            if (threadIdx.x == 0) info[s] = -2;
            __syncthreads();
            continue;
        }
        const DV A{Aall + e.a_offset, 1, n, n * n};
        const DV B{Ball + e.b_offset, 1, n, n * nrhs};
        int64_t st = 0;
        if (e.kind == BATCH_CHOLESKY) {
            st = wg_potrf<N>(A, n, chol_block, piv_s, &flag);
            if (st == 0) {
                wg_systolic<N>(B, A, A, n, nrhs, true, false, xs);                        // L y = b
                wg_backward<N>(B, DV{A.p, A.cs, A.rs, A.ls}, A, n, nrhs, false, exact_backward, xs);  // L' x = y
            }
            if (threadIdx.x == 0) origmax[s * N] = 0.0;
        } else {
            const DV dsub{Dall + e.meta_offset, 1, n, n};
            int64_t *piv = Pall + e.meta_offset;
            uint8_t *blk = Kall + e.meta_offset;
            for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
                piv[i] = i + 1;
                blk[i] = 0;
                stv<N>(dsub, i, 0, mf::zero<N>());
            }
            __syncthreads();
            if (!wg_all_finite_lower<N>(A, n, &flag)) {
                st = -1;
            } else {
                for (int64_t x = threadIdx.x; x < n * n; x += blockDim.x) {  // mirror lower to upper
                    const int64_t r = x % n, c = x / n;
                    if (r > c) stv<N>(A, c, r, ld<N>(A, r, c));
                }
                __syncthreads();
                mf::V<N> mx = mf::zero<N>();
                for (int64_t x = threadIdx.x; x < n * n; x += blockDim.x) {
                    const int64_t r = x % n, c = x / n;
                    if (r >= c) mx = mf::vmax(mx, mf::vabs(ld<N>(A, r, c)));
                }
                mx = wg_max<N>(mx, sv_buf);
                if (threadIdx.x == 0) {
                    for (int l = 0; l < N; ++l) origmax[s * N + l] = mx.x[l];
                    info_s = 0;
                }
                __syncthreads();
                ldlt_unblocked_body<N>(A, n, dsub, piv, blk, alpha, &info_s, sv_buf, si_buf, cbuf);
                __syncthreads();
                st = info_s;
                __syncthreads();
                if (st == 0) {
                    if (threadIdx.x < nrhs) {                                           // forward interchanges
                        const int c = threadIdx.x;
                        for (int64_t k = 0; k < n;) {
                            const int64_t pk = piv[k] - 1;
                            if (blk[k] == 1) { if (pk != k) swap_el<N>(B, k, c, pk, c); k += 1; }
                            else { if (pk != k + 1) swap_el<N>(B, k + 1, c, pk, c); k += 2; }
                        }
                    }
                    __syncthreads();
                    wg_systolic<N>(B, A, A, n, nrhs, true, true, xs);                    // L (unit)
                    for (int64_t x = threadIdx.x; x < n * nrhs; x += blockDim.x) {       // D (1x1 / 2x2)
                        const int64_t k = x % n, c = x / n;
                        if (blk[k] == 1) {
                            stv<N>(B, k, c, mf::div(ld<N>(B, k, c), ld<N>(A, k, k)));
                        } else if (blk[k] == 2) {
                            mf::V<N> s1, s2;
                            solve2x2<N>(ld<N>(A, k, k), ld<N>(dsub, k, 0), ld<N>(A, k + 1, k + 1), ld<N>(B, k, c), ld<N>(B, k + 1, c), s1, s2);
                            stv<N>(B, k, c, s1);
                            stv<N>(B, k + 1, c, s2);
                        }
                    }
                    __syncthreads();
                    wg_backward<N>(B, DV{A.p, A.cs, A.rs, A.ls}, A, n, nrhs, true, exact_backward, xs);  // L' (unit)
                    if (threadIdx.x < nrhs) {                                           // reverse interchanges
                        const int c = threadIdx.x;
                        for (int64_t k = n - 1; k >= 0;) {
                            if (blk[k] == 1) { const int64_t pk = piv[k] - 1; if (pk != k) swap_el<N>(B, k, c, pk, c); k -= 1; }
                            else { const int64_t pk = piv[k - 1] - 1; if (pk != k) swap_el<N>(B, k, c, pk, c); k -= 2; }
                        }
                    }
                    __syncthreads();
                }
            }
        }
        if (threadIdx.x == 0) info[s] = st;
        __syncthreads();
    }
}

template <int N, int BT>
static void launch_shape(hipStream_t s, const BatchArgs &a) {
    S<N> al;
    for (int l = 0; l < N; ++l) al.x[l] = a.alpha[l];
    hipLaunchKernelGGL((k_batched<N, BT>), dim3((unsigned)a.workgroups), dim3(BT), 0, s, a.entries, a.count, a.A, a.B, a.dsub,
                       a.pivots, a.blocks, a.info, a.origmax, a.counter, al, a.chol_block, a.exact_backward);
}

// The launch of one workgroup shape. Each shape is instantiated in a
// translation unit of its own (k_batched.hip, k_batched_64.hip,
// k_batched_128.hip): with several instances of the kernel in one unit the
// compiler stops inlining the shared device routines and the kernel spills.
template <int BT>
static hipError_t launch_bt(int N, hipStream_t s, const BatchArgs &a) {
    switch (N) {
        case 1: launch_shape<1, BT>(s, a); break;
        case 2: launch_shape<2, BT>(s, a); break;
        case 3: launch_shape<3, BT>(s, a); break;
        case 4: launch_shape<4, BT>(s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace mfk
