// Triangular solves for Float64xN (trsm!/trsv!, src/kernels/trsm.jl,
// src/kernels/trsv.jl).
//
// Every solve is expressed on a strided "solution view" X(p, c) (p: unknown
// index, c: right-hand side) and a coefficient view CA(p, q) (the coefficient
// of X(q) in equation p), so left/right and N/T variants are the same kernels
// with different strides. MFLA's element order is
//     v = X(p); for q in substitution order (ascending): v = v - CA(p,q)*X(q);
//     v = v / A(p,p) unless unit
// (trsm.jl:40-50, 63-73, 136-149, 159-172).
//
// Forward substitution (dependencies on q < p) is blocked right-looking:
// a systolic diagonal-block solve (one pivot per step, all later rows of the
// block subtract it) followed by a MODE_SUB tile update of the rows below that
// subtracts the block's terms in ascending q into the running values. Every
// element therefore sees exactly MFLA's ascending-q chain: bitwise equal.
//
// Backward substitution with MFLA's ascending-q chain has an O(n^2) serial
// dependency per right-hand side (row p's first term needs X(p+1), the last
// unknown solved). Two orders are provided: the exact MFLA order
// (k_trsm_seq, one thread per right-hand side) and a blocked right-looking
// order (descending blocks) whose rounding differs from MFLA's but stays
// within the same c*n*u bound.
#include <type_traits>
#include "mf_common.hpp"
#include "k_blas3.hpp"
#include "k_trsm.hpp"
#include "k_factor_dev.hpp"

namespace mfk {

constexpr int TB = 64;   // diagonal block rows (threads in x)
constexpr int TC = 4;    // right-hand sides per workgroup (threads in y)

template <int N, bool FWD>
__global__ __launch_bounds__(TB *TC) void k_trsm_diag(int64_t p0, int nbk, int64_t nrhs, DV X, DV CA, DV D, int unit,
                                                      int swap, const Status *st, int64_t lim) {
    if (st && st->info != 0 && (lim == 0 || st->info < 0 || st->info <= lim)) return;
    __shared__ double xs[TC * N];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int64_t c = (int64_t)blockIdx.x * TC + ty;
    const bool valid = (tx < nbk) && (c < nrhs);
    const int64_t p = p0 + tx;
    mf::V<N> v = valid ? ld<N>(X, p, c) : mf::zero<N>();
    for (int s0 = 0; s0 < nbk; ++s0) {
        const int s = FWD ? s0 : nbk - 1 - s0;
        const int64_t q = p0 + s;
        if (tx == s && valid) {
            if (!unit) v = mf::div(v, ld<N>(D, q, q));
            stv<N>(X, q, c, v);
#pragma unroll
            for (int l = 0; l < N; ++l) xs[ty * N + l] = v.x[l];
        }
        __syncthreads();
        if (valid && (FWD ? tx > s : tx < s)) {
            mf::V<N> xq;
#pragma unroll
            for (int l = 0; l < N; ++l) xq.x[l] = xs[ty * N + l];
            v = swap ? mf::msub(v, xq, ld<N>(CA, p, q)) : mf::msub(v, ld<N>(CA, p, q), xq);
        }
        __syncthreads();
    }
}

// The same diagonal-block solve without barriers: a group of G lanes holds
// one right-hand side (lane p_local = unknown p0 + p_local). At step s the
// owner lane's running value is final; the group reads it by shuffle and
// every lane of the group forms the same quotient x_s (so no lane waits for
// another), then the later unknowns subtract their term. Per element the
// operations and their order are those of k_trsm_diag.
//
// With look-ahead these solves run beside the side stream's tile updates:
// held to 64 VGPRs (8 waves per SIMD; 80 for four limbs, whose previous-block
// terms would spill otherwise) they fit next to the update's waves, and their
// raised wave priority lets the SIMD issue them first.
//
// The serial chain waits on no global load: the block's diagonal is staged
// in LDS once (all groups of the workgroup share the block), and each step's
// coefficient CA(p, q) is loaded before the step's division -- one step
// ahead for up to three limbs; at the top of the step for four, where a
// second coefficient in registers would spill under the 64-VGPR bound. The
// loads use a row index clamped into the block (lanes that do not use the
// value read an in-bounds entry they ignore).
// Start mark for the side stream: *start_sig = start_value from thread 0 of
// every workgroup (the same value; any one suffices). A kernel starts only
// after the kernel before it on its stream has ended and released its
// writes, so the mark says those writes are done; the side stream's wait
// kernel acquires before the update that reads them.
__device__ inline void mark_start(unsigned *start_sig, unsigned start_value) {
    if (start_sig && threadIdx.x == 0) __hip_atomic_store(start_sig, start_value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int G>
__device__ __forceinline__ double group_bcast(double x, int grp, int s) {
    if constexpr (G == 64) {
        // one group per wave: the source lane is wave-uniform
        const int64_t b = __builtin_bit_cast(int64_t, x);
        const int lo = __builtin_amdgcn_readlane((int)b, s), hi = __builtin_amdgcn_readlane((int)(b >> 32), s);
        return __builtin_bit_cast(double, ((int64_t)hi << 32) | (uint32_t)lo);
    } else {
        return __shfl(x, grp * G + s, 64);
    }
}

//
// With K > 0 the kernel first subtracts the K terms of the block solved just
// before (unknowns q0 .. q0 + K - 1, ascending q: the order and operations of
// the MODE_SUB tile update it replaces), so that update is not a kernel of
// its own on the serial chain of blocks.
//
// With far_flag the kernel first waits (bounded) for *far_flag >= far_wait:
// the side stream's update that last reaches this block's rows was enqueued
// before this kernel and waits only for kernels before it, so the wait ends;
// past the bound the block's values are written as NaN and *err = 1.
template <int N, bool FWD, int G>
__global__ __launch_bounds__(256, N == 4 ? 5 : 8) void k_trsm_diag_wave(int64_t p0, int nbk, int64_t nrhs, DV X, DV CA, DV D, int unit, int swap,
                                                           const Status *st, int64_t lim, int64_t q0, int K, const unsigned *far_flag,
                                                           unsigned far_wait, unsigned *err, unsigned *start_sig, unsigned start_value) {
    __builtin_amdgcn_s_setprio(3);
    mark_start(start_sig, start_value);
    if (st && st->info != 0 && (lim == 0 || st->info < 0 || st->info <= lim)) return;
    const bool ok = far_flag == nullptr || wait_flag(far_flag, far_wait);
    __shared__ double dg[G * N];
    constexpr int RPW = 64 / G;
    const int lane = threadIdx.x & 63, grp = lane / G, pl = lane % G;
    const int64_t c = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * RPW + grp;
    const bool valid = pl < nbk && c < nrhs;
    const int64_t p = p0 + pl;
    const int64_t pc = p0 + (pl < nbk ? pl : nbk - 1);   // in-bounds row for the unconditional loads
    if (!unit && (int)threadIdx.x < nbk) lds_put<N>(&dg[threadIdx.x * N], ld<N>(D, p0 + threadIdx.x, p0 + threadIdx.x));
    constexpr bool kAhead = N <= 3;
    mf::V<N> v = valid ? ld<N>(X, p, c) : mf::zero<N>();
    if (!ok) {
#pragma unroll
        for (int l = 0; l < N; ++l) v.x[l] = __builtin_nan("");
        if (threadIdx.x == 0) __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (K > 0) {
        // terms of the previous block (every lane computes; invalid lanes'
        // values are never stored), the next term's operands loaded ahead
        const int64_t cc = c < nrhs ? c : nrhs - 1;
        mf::V<N> an2 = ld<N>(CA, pc, q0), xn2 = ld<N>(X, q0, cc);
        for (int k = 0; k < K; ++k) {
            const mf::V<N> a = an2, xq = xn2;
            if (k + 1 < K) {
                an2 = ld<N>(CA, pc, q0 + k + 1);
                xn2 = ld<N>(X, q0 + k + 1, cc);
            }
            v = swap ? mf::msub(v, xq, a) : mf::msub(v, a, xq);
        }
    }
    mf::V<N> an = kAhead ? ld<N>(CA, pc, p0 + (FWD ? 0 : nbk - 1)) : mf::zero<N>();
    __syncthreads();
    for (int s0 = 0; s0 < nbk; ++s0) {
        const int s = FWD ? s0 : nbk - 1 - s0;
        const int64_t q = p0 + s;
        mf::V<N> a;
        if constexpr (kAhead) {
            a = an;
            if (s0 + 1 < nbk) an = ld<N>(CA, pc, FWD ? q + 1 : q - 1);
        } else {
            a = ld<N>(CA, pc, q);
        }
        mf::V<N> xq;
#pragma unroll
        for (int l = 0; l < N; ++l) xq.x[l] = group_bcast<G>(v.x[l], grp, s);
        if (!unit) xq = mf::div(xq, lds_get<N>(&dg[s * N]));
        if (valid) {
            if (pl == s) {
                v = xq;
                stv<N>(X, q, c, xq);
            } else if (FWD ? pl > s : pl < s) {
                v = swap ? mf::msub(v, xq, a) : mf::msub(v, a, xq);
            }
        }
    }
}

// Forward look-ahead with the next block's rows streamed. A workgroup of 8
// waves takes 4 right-hand sides: waves 0-3 solve the block (the operations
// of k_trsm_diag_wave with one wave per right-hand side and no previous-block
// terms); waves 4-7 take the next block's rows p1 + lane and, once the side
// stream's update from the block before has reached them (*far_flag >=
// far_wait, written after that update), subtract this block's terms in
// ascending q as the solver waves publish them in LDS -- the operands,
// operations and order of the next kernel's previous-block terms
// (k_trsm_diag_wave with K > 0), which then has none. That update was
// enqueued before this kernel and waits only for the kernel before it, so
// the wait ends; it is bounded all the same, and past the bound the rows
// are written as NaN (and *err = 1) instead of values without the update.
constexpr int kStreamRhs = 4;
template <int N>
__global__ __launch_bounds__(512, N == 4 ? 3 : 4) void k_trsm_diag_stream(int64_t p0, int nbk, int64_t p1, int nn, int64_t nrhs, DV X,
                                                                        DV CA, DV D, int unit, int swap, const Status *st, int64_t lim,
                                                                        const unsigned *far_flag, unsigned far_wait, unsigned *err,
                                                                        unsigned *start_sig, unsigned start_value) {
    mark_start(start_sig, start_value);
    if (st && st->info != 0 && (lim == 0 || st->info < 0 || st->info <= lim)) return;
    __shared__ double dg[64 * N];
    __shared__ double xs[kStreamRhs * 64 * N];
    __shared__ int cnt[kStreamRhs];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const bool streamer = wave >= kStreamRhs;
    const int r = streamer ? wave - kStreamRhs : wave;
    const int64_t c = (int64_t)blockIdx.x * kStreamRhs + r;
    const int64_t cc = c < nrhs ? c : nrhs - 1;   // in-bounds column for the unconditional loads
    if (!unit && (int)threadIdx.x < nbk) lds_put<N>(&dg[threadIdx.x * N], ld<N>(D, p0 + threadIdx.x, p0 + threadIdx.x));
    if ((int)threadIdx.x < kStreamRhs) cnt[threadIdx.x] = 0;
    __syncthreads();
    if (!streamer) {
        __builtin_amdgcn_s_setprio(3);
        const bool valid = lane < nbk && c < nrhs;
        const int64_t p = p0 + lane;
        const int64_t pc = p0 + (lane < nbk ? lane : nbk - 1);
        constexpr bool kAhead = N <= 3;
        mf::V<N> v = valid ? ld<N>(X, p, c) : mf::zero<N>();
        mf::V<N> an = kAhead ? ld<N>(CA, pc, p0) : mf::zero<N>();
        for (int s = 0; s < nbk; ++s) {
            const int64_t q = p0 + s;
            mf::V<N> a;
            if constexpr (kAhead) {
                a = an;
                if (s + 1 < nbk) an = ld<N>(CA, pc, q + 1);
            } else {
                a = ld<N>(CA, pc, q);
            }
            mf::V<N> xq;
#pragma unroll
            for (int l = 0; l < N; ++l) xq.x[l] = group_bcast<64>(v.x[l], 0, s);
            if (!unit) xq = mf::div(xq, lds_get<N>(&dg[s * N]));
            if (lane == 0) {
                lds_put<N>(&xs[(r * 64 + s) * N], xq);
                __hip_atomic_store(&cnt[r], s + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            if (valid) {
                if (lane == s) {
                    v = xq;
                    stv<N>(X, q, c, xq);
                } else if (lane > s) {
                    v = swap ? mf::msub(v, xq, a) : mf::msub(v, a, xq);
                }
            }
        }
    } else {
        const bool valid = lane < nn && c < nrhs;
        const int64_t p = p1 + lane, pc = p1 + (lane < nn ? lane : nn - 1);
        const bool ok = wait_flag(far_flag, far_wait);
        mf::V<N> v = ld<N>(X, pc, cc);
        mf::V<N> an = ld<N>(CA, pc, p0);
        for (int s = 0; s < nbk; ++s) {
            const mf::V<N> a = an;
            if (s + 1 < nbk) an = ld<N>(CA, pc, p0 + s + 1);
            while (__hip_atomic_load(&cnt[r], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) <= s) __builtin_amdgcn_s_sleep(1);
            const mf::V<N> xq = lds_get<N>(&xs[(r * 64 + s) * N]);
            v = swap ? mf::msub(v, xq, a) : mf::msub(v, a, xq);
        }
        if (!ok) {
#pragma unroll
            for (int l = 0; l < N; ++l) v.x[l] = __builtin_nan("");
            if (lane == 0) __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (valid) stv<N>(X, p, c, v);
    }
}

// *flag = value after the work before it on the stream
__global__ void k_trsm_flag(unsigned *flag, unsigned value) {
    if (threadIdx.x != 0) return;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_store(flag, value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

hipError_t flag_launch(hipStream_t s, unsigned *flag, unsigned value) {
    hipLaunchKernelGGL(k_trsm_flag, dim3(1), dim3(64), 0, s, flag, value);
    return hipGetLastError();
}

// On the side stream before a block's far update: waits (bounded) for the
// start mark of the diagonal kernel after that block's (*start_sig >=
// value). Past the bound it sets *err = 1 and writes NaN over the rows the
// update writes, so the failure shows in the solution.
template <int N>
__global__ __launch_bounds__(256) void k_trsm_side_wait(const unsigned *start_sig, unsigned value, unsigned *err, DV X, int64_t r0,
                                                        int64_t rows, int64_t nrhs) {
    __shared__ int ok;
    if (threadIdx.x == 0) ok = wait_flag(start_sig, value) ? 1 : 0;
    __syncthreads();
    if (ok) return;
    if (threadIdx.x == 0) __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    mf::V<N> bad;
#pragma unroll
    for (int l = 0; l < N; ++l) bad.x[l] = __builtin_nan("");
    for (int64_t e = threadIdx.x; e < rows * nrhs; e += blockDim.x) stv<N>(X, r0 + e % rows, e / rows, bad);
}

// Exact MFLA order for backward substitution: one thread per right-hand side.
template <int N>
__global__ void k_trsm_seq_bwd(int64_t n, int64_t nrhs, DV X, DV CA, DV D, int unit, int swap, const Status *st, int64_t lim) {
    if (st && st->info != 0 && (lim == 0 || st->info < 0 || st->info <= lim)) return;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nrhs) return;
    for (int64_t p = n - 1; p >= 0; --p) {
        mf::V<N> v = ld<N>(X, p, c);
        for (int64_t q = p + 1; q < n; ++q)
            v = swap ? mf::msub(v, ld<N>(X, q, c), ld<N>(CA, p, q)) : mf::msub(v, ld<N>(CA, p, q), ld<N>(X, q, c));
        if (!unit) v = mf::div(v, ld<N>(D, p, p));
        stv<N>(X, p, c, v);
    }
}

static DV sub(const DV &v, int64_t i, int64_t j) { return DV{v.p + i * v.rs + j * v.cs, v.rs, v.cs, v.ls}; }

template <int N>
static hipError_t diag(hipStream_t s, bool fwd, int64_t p0, int nbk, int64_t nrhs, const DV &X, const DV &CA, const DV &D, int unit,
                       int swap, const Status *st, int64_t lim, int64_t q0 = 0, int K = 0, const unsigned *far_flag = nullptr,
                       unsigned far_wait = 0, unsigned *err = nullptr, unsigned *start_sig = nullptr, unsigned start_value = 0) {
    // one group of G lanes per right-hand side, 4 waves per workgroup
    auto launch = [&](auto gtag) {
        constexpr int G = decltype(gtag)::value;
        const dim3 grid((unsigned)cdiv(cdiv(nrhs, 64 / G), 4)), block(256);
        if (fwd)
            hipLaunchKernelGGL((k_trsm_diag_wave<N, true, G>), grid, block, 0, s, p0, nbk, nrhs, X, CA, D, unit, swap, st, lim, q0, K, far_flag,
                               far_wait, err, start_sig, start_value);
        else
            hipLaunchKernelGGL((k_trsm_diag_wave<N, false, G>), grid, block, 0, s, p0, nbk, nrhs, X, CA, D, unit, swap, st, lim, q0, K, far_flag,
                               far_wait, err, start_sig, start_value);
    };
    if ((far_flag || start_sig) && nbk > TB) return hipErrorInvalidValue;
    if (nbk <= 16) launch(std::integral_constant<int, 16>());
    else if (nbk <= 32) launch(std::integral_constant<int, 32>());
    else if (nbk <= TB) launch(std::integral_constant<int, 64>());
    else {
        dim3 grid((unsigned)cdiv(nrhs, TC)), block(TB, TC);
        if (fwd) hipLaunchKernelGGL((k_trsm_diag<N, true>), grid, block, 0, s, p0, nbk, nrhs, X, CA, D, unit, swap, st, lim);
        else hipLaunchKernelGGL((k_trsm_diag<N, false>), grid, block, 0, s, p0, nbk, nrhs, X, CA, D, unit, swap, st, lim);
    }
    return hipGetLastError();
}

template <int N>
static hipError_t solve_t(hipStream_t s, const TrsmArgs &a, const DV &X, const DV &CA, const DV &D, const Status *st) {
    const int64_t n = a.n, nrhs = a.nrhs;
    if (n == 0 || nrhs == 0) return hipSuccess;
    hipError_t e = hipSuccess;
    // The updates take 16 x 16 tiles: with K = 64 a 32 x 32 tile's grid
    // (e.g. 992 workgroups for 3968 rows x 256 right-hand sides, at three
    // per CU) needs a second, mostly idle round, while four times as many
    // 16 x 16 workgroups fill ~4 rounds (x4 4096 x 256: 22.3 -> 20.0 ms; x2 /
    // x3 and 1024 right-hand sides faster too). The tile does not change
    // the bits.
    TileArgs t{};
    t.mode = MODE_SUB; t.net = mf::NET_STD; t.lower = 0; t.beta_zero = 0; t.tile = 0; t.row_lo = 0; t.n = nrhs;
    t.swap_product = a.swap_product;
    t.status_limit = a.status_limit;
    // X(r0 : r0 + rows) -= CA(r0 : , q0 : q0 + K) X(q0 : q0 + K) on stream ss
    auto update = [&](hipStream_t ss, int64_t r0, int64_t rows, int64_t q0, int64_t K) -> hipError_t {
        if (rows <= 0) return hipSuccess;
        TileArgs u = t;
        u.rows = rows; u.K = K;
        return tile_launch_n(N, ss, u, sub(CA, r0, q0), sub(X, q0, 0), sub(X, r0, 0), nullptr, nullptr, st);
    };
    // Look-ahead (blocked orders, n > 3 blocks, a side stream): the terms of
    // block b reach the rows of block b + 1 in the next block's diagonal
    // kernel (its prologue), and the rows beyond on the side stream while
    // that kernel runs. Block b's kernel therefore waits for the side
    // stream's update from block b - 2 (the last one covering its rows);
    // every element still takes the blocks' terms in MFLA's block order, and
    // each block's terms in ascending q.
    const bool ahead = a.side != nullptr && n > 3 * TB;
    if (ahead && a.far_flag) {
        // With device counters (far_flag, start_sig): the side stream writes
        // *far_flag = b + 1 after block b's far update, and block b's kernel
        // waits on it (for the update that last reaches its rows) instead of
        // the stream waiting on an event (~11 us each even when satisfied).
        // Block b's far update waits on the side stream for the start mark
        // of block b + 1's kernel (*start_sig >= b + 1: block b is solved)
        // instead of an event recorded between the two diagonal kernels
        // (~6.7 us each on the chain); the host enqueues it after block
        // b + 1's kernel. Then every counter a kernel waits for is advanced by
        // work enqueued before it, except in the streamed order below, where
        // block b + 1's kernel waits for block b's far update: that order
        // defers the far updates only between streams with hardware queues
        // of their own (dedicated), and otherwise enqueues each far update
        // after an event recorded behind its block's kernel.
        if (!a.start_sig) return hipErrorInvalidValue;
        hipError_t r;
        if ((r = flag_launch(s, a.far_flag, 0u)) != hipSuccess) return r;
        if ((r = flag_launch(s, a.start_sig, 0u)) != hipSuccess) return r;
        // the side stream after the resets (it waited for the caller's work)
        if ((r = hipEventRecord(a.ev_main, s)) != hipSuccess) return r;
        if ((r = hipStreamWaitEvent(a.side, a.ev_main, 0)) != hipSuccess) return r;
        struct Far { int64_t r0, rows, q0; int K; unsigned b; };
        Far pend{0, 0, 0, 0, 0};
        bool any_far = false;
        auto far_update = [&](const Far &f) -> hipError_t {
            hipError_t q;
            hipLaunchKernelGGL(k_trsm_side_wait<N>, dim3(1), dim3(256), 0, a.side, a.start_sig, f.b + 1, a.far_flag + 1, X, f.r0, f.rows, nrhs);
            if ((q = hipGetLastError()) != hipSuccess) return q;
            if ((q = update(a.side, f.r0, f.rows, f.q0, f.K)) != hipSuccess) return q;
            any_far = true;
            return flag_launch(a.side, a.far_flag, f.b + 1);
        };
        unsigned b = 0;
        bool had_far[2] = {false, false};    // blocks b - 2, b - 1 (by parity)
        if (a.forward && a.stream_next) {
            // The next block's rows streamed (k_trsm_diag_stream): block b's
            // kernel applies its terms to block b + 1's rows once the far
            // update from block b - 1 has reached them (*far_flag >= b); no
            // diagonal kernel has previous-block terms of its own.
            for (int64_t p0 = 0; p0 < n; p0 += TB, ++b) {
                const int nbk = (int)((n - p0) < TB ? (n - p0) : TB);
                const int64_t p1 = p0 + nbk;
                const int nn = (int)((n - p1) < TB ? (n - p1) : TB);
                if (nn > 0) {
                    hipLaunchKernelGGL(k_trsm_diag_stream<N>, dim3((unsigned)cdiv(nrhs, kStreamRhs)), dim3(512), 0, s, p0, nbk, p1, nn, nrhs, X,
                                       CA, D, a.unit, a.swap_product, st, a.status_limit, a.far_flag, b, a.far_flag + 1, a.start_sig, b);
                    if ((r = hipGetLastError()) != hipSuccess) return r;
                } else if ((r = diag<N>(s, true, p0, nbk, nrhs, X, CA, D, a.unit, a.swap_product, st, a.status_limit, 0, 0, nullptr, 0, nullptr,
                                        a.start_sig, b)) != hipSuccess) {
                    return r;
                }
                if (pend.rows > 0) {
                    if ((r = far_update(pend)) != hipSuccess) return r;
                    pend.rows = 0;
                }
                const int64_t far0 = p1 + nn;
                if (far0 < n) {
                    pend = Far{far0, n - far0, p0, nbk, b};
                    if (!a.dedicated) {
                        // block b solved: the start mark is not awaited
                        if ((r = flag_launch(s, a.start_sig, b + 1)) != hipSuccess) return r;
                        if ((r = far_update(pend)) != hipSuccess) return r;
                        pend.rows = 0;
                    }
                }
            }
        } else {
            // Block b's kernel takes block b - 1's terms for its rows in its
            // prologue (K > 0) and waits for block b - 2's far update.
            int64_t prev0 = 0;
            int prevK = 0;
            auto block_step = [&](int64_t p0, int nbk, int64_t far0, int64_t nfar) -> hipError_t {
                hipError_t q;
                const unsigned wait = (b >= 2 && had_far[b & 1]) ? b - 1 : 0;
                if ((q = diag<N>(s, a.forward != 0, p0, nbk, nrhs, X, CA, D, a.unit, a.swap_product, st, a.status_limit, prev0, prevK,
                                 wait ? a.far_flag : nullptr, wait, a.far_flag + 1, a.start_sig, b)) != hipSuccess)
                    return q;
                if (pend.rows > 0) {
                    if ((q = far_update(pend)) != hipSuccess) return q;
                    pend.rows = 0;
                }
                had_far[b & 1] = nfar > 0;
                if (nfar > 0) pend = Far{far0, nfar, p0, nbk, b};
                prev0 = p0;
                prevK = nbk;
                ++b;
                return hipSuccess;
            };
            if (a.forward) {
                for (int64_t p0 = 0; p0 < n; p0 += TB) {
                    const int nbk = (int)((n - p0) < TB ? (n - p0) : TB);
                    const int64_t p1 = p0 + nbk, nn = (n - p1) < TB ? (n - p1) : TB;
                    if ((r = block_step(p0, nbk, p1 + nn, n - p1 - nn)) != hipSuccess) return r;
                }
            } else {
                for (int64_t p1 = n; p1 > 0; p1 -= TB) {
                    const int64_t p0 = p1 > TB ? p1 - TB : 0;
                    const int64_t nn = p0 < TB ? p0 : TB;
                    if ((r = block_step(p0, (int)(p1 - p0), 0, p0 - nn)) != hipSuccess) return r;
                }
            }
        }
        if (pend.rows > 0) {
            // no kernel after the last one with a far update: mark it here
            if ((r = flag_launch(s, a.start_sig, b)) != hipSuccess) return r;
            if ((r = far_update(pend)) != hipSuccess) return r;
        }
        if (any_far) {
            if ((r = hipEventRecord(a.ev_side, a.side)) != hipSuccess) return r;
            if ((r = hipStreamWaitEvent(s, a.ev_side, 0)) != hipSuccess) return r;
        }
        return hipSuccess;
    }
    if (ahead) {
        // Event waits: block b's kernel after block b - 2's far update on the
        // side stream, each far update after its block's kernel.
        hipEvent_t far_ev[2] = {a.ev_side, a.ev_side2};
        bool far_pending[2] = {false, false};
        int64_t prev0 = 0;
        int prevK = 0;
        int64_t b = 0;
        auto block_step = [&](int64_t p0, int nbk, int64_t far0, int64_t nfar) -> hipError_t {
            hipError_t r;
            const int slot = (int)(b & 1);
            if (far_pending[slot]) {   // the update from block b - 2
                if ((r = hipStreamWaitEvent(s, far_ev[slot], 0)) != hipSuccess) return r;
                far_pending[slot] = false;
            }
            if ((r = diag<N>(s, a.forward != 0, p0, nbk, nrhs, X, CA, D, a.unit, a.swap_product, st, a.status_limit, prev0, prevK)) !=
                hipSuccess)
                return r;
            if (nfar > 0) {
                if ((r = hipEventRecord(a.ev_main, s)) != hipSuccess) return r;
                if ((r = hipStreamWaitEvent(a.side, a.ev_main, 0)) != hipSuccess) return r;
                if ((r = update(a.side, far0, nfar, p0, nbk)) != hipSuccess) return r;
                if ((r = hipEventRecord(far_ev[slot], a.side)) != hipSuccess) return r;
                far_pending[slot] = true;
            }
            prev0 = p0;
            prevK = nbk;
            ++b;
            return hipSuccess;
        };
        if (a.forward) {
            for (int64_t p0 = 0; p0 < n; p0 += TB) {
                const int nbk = (int)((n - p0) < TB ? (n - p0) : TB);
                const int64_t p1 = p0 + nbk, nn = (n - p1) < TB ? (n - p1) : TB;
                if ((e = block_step(p0, nbk, p1 + nn, n - p1 - nn)) != hipSuccess) return e;
            }
        } else {
            for (int64_t p1 = n; p1 > 0; p1 -= TB) {
                const int64_t p0 = p1 > TB ? p1 - TB : 0;
                const int64_t nn = p0 < TB ? p0 : TB;
                if ((e = block_step(p0, (int)(p1 - p0), 0, p0 - nn)) != hipSuccess) return e;
            }
        }
        for (int q = 0; q < 2; ++q)
            if (far_pending[q] && (e = hipStreamWaitEvent(s, far_ev[q], 0)) != hipSuccess) return e;
        return hipSuccess;
    }
    // Without look-ahead: each block's update of the rows below (above) it
    // right after its solve, on the caller's stream.
    if (a.forward) {
        for (int64_t p0 = 0; p0 < n; p0 += TB) {
            const int nbk = (int)((n - p0) < TB ? (n - p0) : TB);
            if ((e = diag<N>(s, true, p0, nbk, nrhs, X, CA, D, a.unit, a.swap_product, st, a.status_limit)) != hipSuccess) return e;
            const int64_t p1 = p0 + nbk;
            if (p1 < n && (e = update(s, p1, n - p1, p0, nbk)) != hipSuccess) return e;
        }
        return hipSuccess;
    }
    if (a.exact_order) {
        hipLaunchKernelGGL(k_trsm_seq_bwd<N>, dim3((unsigned)cdiv(nrhs, 64)), dim3(64), 0, s, n, nrhs, X, CA, D, a.unit, a.swap_product, st, a.status_limit);
        return hipGetLastError();
    }
    for (int64_t p1 = n; p1 > 0; p1 -= TB) {
        const int64_t p0 = p1 > TB ? p1 - TB : 0;
        const int nbk = (int)(p1 - p0);
        if ((e = diag<N>(s, false, p0, nbk, nrhs, X, CA, D, a.unit, a.swap_product, st, a.status_limit)) != hipSuccess) return e;
        if (p0 > 0 && (e = update(s, 0, p0, p0, nbk)) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t trsm_solve(int N, hipStream_t s, const TrsmArgs &a, const DV &X, const DV &CA, const DV &D, const Status *st) {
    switch (N) {
        case 1: return solve_t<1>(s, a, X, CA, D, st);
        case 2: return solve_t<2>(s, a, X, CA, D, st);
        case 3: return solve_t<3>(s, a, X, CA, D, st);
        case 4: return solve_t<4>(s, a, X, CA, D, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace mfk
