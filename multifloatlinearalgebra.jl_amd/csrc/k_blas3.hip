// GEMM-family tile kernel for Float64xN on CDNA4.
//
// One kernel body serves gemm!, syrk!, gemmt!, the trailing updates of the
// factorizations and the running-value updates of the triangular solves:
//
//   MODE_AB : acc = 0; for k ascending acc = acc + a(i,k)*b(k,j);
//             C = alpha*acc            (beta == 0, C never read)
//             C = alpha*acc + beta*C   (otherwise)
//             -- src/kernels/gemm.jl:151-181, syrk.jl:13-38, gemmt.jl:15-40
//   MODE_SUB: v = C; for k ascending v = v - a(i,k)*b(k,j); C = v
//             -- the running-value order of src/kernels/trsm.jl:40-47
//
// Each output element is owned by exactly one thread and reduced in ascending
// k with one fixed network, so results are deterministic and independent of
// the tiling. There is no split-K anywhere.
//
// Layout: a 256-thread workgroup computes a (16*TM) x (16*TN) tile; A and B
// k-panels of depth BK are staged through LDS. B is [col][k][limb] with a
// padded column stride (a wave reads 4 distinct column groups: broadcasts);
// A is stored in 16-byte pieces so the 16 row-threads of a wave read 256
// contiguous bytes per piece. Operands are
// strided views, so transposed operands (syrk's P^T P, gemmt's L R^T) need no
// copies.
#include "mf_common.hpp"
#include "k_blas3.hpp"
#include <type_traits>

namespace mfk {

template <int N, int NET, int TM, int TN, int BK, int MODE, bool LOWER>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_tile(int64_t n, int64_t K, DV A, DV B, DV C, S<N> alpha, S<N> beta,
                                              int beta_zero, int64_t row_lo, int64_t row_hi, const Status *status,
                                              int64_t status_limit, int swap) {
    if (status && status_skips(status, status_limit)) return;
    constexpr int BM = 16 * TM, BN = 16 * TN;
    __shared__ double As[BK * BM * N];
    // B's LDS slot of limb l of B(kk, jj). With two or more limbs: column-
    // major [col][k][limb] with the column stride padded by two doubles, so
    // the 64 lanes writing 16 k x 4 columns of a panel spread over the banks
    // and the four column groups a wave reads per k start 8 (mod 32) banks
    // apart; with one limb [k][col] (a thread's columns are adjacent: one
    // 16-byte read).
    constexpr int BCS = BK * N + 2;
    auto b_slot = [](int kk, int jj) -> int {
        if constexpr (N >= 2) return jj * BCS + kk * N;
        else return kk * BN + jj;
    };
    __shared__ double Bs[N >= 2 ? BN * BCS : BK * BN];

    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    // LDS slot of limb l of A(ii, kk). With TM * N even, the TM * N doubles a
    // thread reads per k are stored as 16-byte pieces, piece-major over the
    // 16 row-threads: the 16 distinct threads of a wave read one piece as 256
    // contiguous bytes (no bank conflicts); otherwise [k][row][limb].
    auto a_slot = [](int kk, int ii, int l) -> int {
        if constexpr ((TM * N) % 2 == 0) {
            constexpr int P = TM * N / 2;
            const int owner = ii / TM, d = (ii % TM) * N + l;
            return ((kk * P + (d >> 1)) * 16 + owner) * 2 + (d & 1);
        } else {
            return (kk * BM + ii) * N + l;
        }
    };
    const int64_t i0 = row_lo + (int64_t)blockIdx.x * BM;
    const int64_t j0 = (int64_t)blockIdx.y * BN;
    if (LOWER && i0 + BM - 1 < j0) return;  // tile strictly above the diagonal

    mf::V<N> acc[TM][TN];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) {
            if (MODE == MODE_SUB) {
                const int64_t i = i0 + tx * TM + a, j = j0 + ty * TN + b;
                acc[a][b] = (i < row_hi && j < n) ? ld<N>(C, i, j) : mf::zero<N>();
            } else {
                acc[a][b] = mf::zero<N>();
            }
        }

    const bool a_rowfast = (A.rs == 1);
    const bool b_kfast = (B.rs == 1);
    // Panels wholly inside the operands (every panel of an interior tile but
    // a partial last one) are staged without bounds checks: per limb a
    // uniform panel base, advanced by BK per panel, plus a per-thread 32-bit
    // element offset fixed for the whole reduction; the LDS slots are fixed
    // too. Element q of a thread (e = tid + 256 q) differs from element 0 by a
    // uniform step in k (or in rows / columns), so one offset and one slot per
    // operand are kept. The elements and slots are those of the checked path.
    constexpr int AE = BM * BK / 256, BE = BN * BK / 256;
    // (the 32-bit offsets need the panels' extents below 2^31 elements)
    // Kept to three or four limbs and the smaller tiles of one or two: with
    // the one- and two-limb 32x64 / 64x64 tiles the second staging path
    // costs registers the occupancy needs.
    constexpr bool kFastStage = N >= 3 || TM * TN <= 4;
    const bool interior = kFastStage && i0 + BM <= row_hi && j0 + BN <= n && (BM + BK) * (A.rs + A.cs) < INT32_MAX &&
                          (BN + BK) * (B.rs + B.cs) < INT32_MAX;
    const int a_ii = a_rowfast ? tid % BM : tid / BK, a_kk = a_rowfast ? tid / BM : tid % BK;
    const int b_jj = b_kfast ? tid / BK : tid % BN, b_kk = b_kfast ? tid % BK : tid / BN;
    // per element q: A (ii, kk + q * 256 / BM) or (ii + q * 256 / BK, kk);
    // B (kk, jj + q * 256 / BK) or (kk + q * 256 / BN, jj)
    const int aoff = (int)(a_ii * A.rs + a_kk * A.cs), boff = (int)(b_kk * B.rs + b_jj * B.cs);
    const int a_step = a_rowfast ? (int)((256 / BM) * A.cs) : (int)((256 / BK) * A.rs);
    const int b_step = b_kfast ? (int)((256 / BK) * B.cs) : (int)((256 / BN) * B.rs);
    auto a_slot_q = [&](int q, int l) {
        return a_rowfast ? a_slot(a_kk + q * (256 / BM), a_ii, l) : a_slot(a_kk, a_ii + q * (256 / BK), l);
    };
    auto b_slot_q = [&](int q) { return b_kfast ? b_slot(b_kk, b_jj + q * (256 / BK)) : b_slot(b_kk + q * (256 / BN), b_jj); };
    const double *abase = A.p + i0 * A.rs, *bbase = B.p + j0 * B.cs;
    // 16 x 16 tiles: the next panel's operands are loaded into registers
    // while the current panel computes (with a short reduction, e.g. K = 64,
    // the workgroups stay in step and each of the few panels would otherwise
    // wait for its loads): x4 potrs 64 RHS 29.0 -> 28.1 ms, trsm 4096 x 256
    // 20.0 -> 19.8 ms; x2 / x3 gain as well despite fewer waves per SIMD
    constexpr bool kPrefetch = kFastStage && TM * TN == 1;
    double pa_[AE][N], pb_[BE][N];
    bool have = false;                              // pa_ / pb_ hold this panel's operands
    for (int64_t k0 = 0; k0 < K; k0 += BK, abase += BK * A.cs, bbase += BK * B.rs) {
        if (kPrefetch && have) {
#pragma unroll
            for (int q = 0; q < AE; ++q)
#pragma unroll
                for (int l = 0; l < N; ++l) As[a_slot_q(q, l)] = pa_[q][l];
#pragma unroll
            for (int q = 0; q < BE; ++q)
#pragma unroll
                for (int l = 0; l < N; ++l) Bs[b_slot_q(q) + l] = pb_[q][l];
        } else if (kFastStage && interior && k0 + BK <= K) {
            // all loads in flight together, unless the accumulators leave no
            // room for them (then A's, then B's)
            constexpr bool kSplit = TM * TN * N + (AE + BE) * N > 32;
            double av_[AE][N], bv_[BE][N];
#pragma unroll
            for (int q = 0; q < AE; ++q)
#pragma unroll
                for (int l = 0; l < N; ++l) av_[q][l] = (abase + l * A.ls)[aoff + q * a_step];
            if constexpr (kSplit) {
#pragma unroll
                for (int q = 0; q < AE; ++q)
#pragma unroll
                    for (int l = 0; l < N; ++l) As[a_slot_q(q, l)] = av_[q][l];
            }
#pragma unroll
            for (int q = 0; q < BE; ++q)
#pragma unroll
                for (int l = 0; l < N; ++l) bv_[q][l] = (bbase + l * B.ls)[boff + q * b_step];
            if constexpr (!kSplit) {
#pragma unroll
                for (int q = 0; q < AE; ++q)
#pragma unroll
                    for (int l = 0; l < N; ++l) As[a_slot_q(q, l)] = av_[q][l];
            }
#pragma unroll
            for (int q = 0; q < BE; ++q)
#pragma unroll
                for (int l = 0; l < N; ++l) Bs[b_slot_q(q) + l] = bv_[q][l];
        } else {
        // stage A panel: BM x BK elements
        for (int e = tid; e < BM * BK; e += 256) {
            int ii, kk;
            if (a_rowfast) { ii = e % BM; kk = e / BM; } else { kk = e % BK; ii = e / BK; }
            const int64_t gi = i0 + ii, gk = k0 + kk;
            const bool ok = gi < row_hi && gk < K;
            const double *src = A.p + gi * A.rs + gk * A.cs;
#pragma unroll
            for (int l = 0; l < N; ++l) As[a_slot(kk, ii, l)] = ok ? src[l * A.ls] : 0.0;
        }
        // stage B panel: BK x BN elements
        for (int e = tid; e < BN * BK; e += 256) {
            int jj, kk;
            if (b_kfast) { kk = e % BK; jj = e / BK; } else { jj = e % BN; kk = e / BN; }
            const int64_t gj = j0 + jj, gk = k0 + kk;
            double *dst = &Bs[b_slot(kk, jj)];
            if (gj < n && gk < K) {
                const double *src = B.p + gk * B.rs + gj * B.cs;
#pragma unroll
                for (int l = 0; l < N; ++l) dst[l] = src[l * B.ls];
            } else {
#pragma unroll
                for (int l = 0; l < N; ++l) dst[l] = 0.0;
            }
        }
        }
        __syncthreads();
        if constexpr (kPrefetch) {
            have = interior && k0 + 2 * BK <= K;
            if (have) {
#pragma unroll
                for (int q = 0; q < AE; ++q)
#pragma unroll
                    for (int l = 0; l < N; ++l) pa_[q][l] = (abase + BK * A.cs + l * A.ls)[aoff + q * a_step];
#pragma unroll
                for (int q = 0; q < BE; ++q)
#pragma unroll
                    for (int l = 0; l < N; ++l) pb_[q][l] = (bbase + BK * B.rs + l * B.ls)[boff + q * b_step];
            }
        }
        const int kend = (K - k0) < BK ? (int)(K - k0) : BK;
        for (int kk = 0; kk < kend; ++kk) {
            mf::V<N> av[TM], bv[TN];
#pragma unroll
            for (int a = 0; a < TM; ++a)
#pragma unroll
                for (int l = 0; l < N; ++l) av[a].x[l] = As[a_slot(kk, tx * TM + a, l)];
#pragma unroll
            for (int b = 0; b < TN; ++b)
#pragma unroll
                for (int l = 0; l < N; ++l) bv[b].x[l] = Bs[b_slot(kk, ty * TN + b) + l];
#pragma unroll
            for (int a = 0; a < TM; ++a)
#pragma unroll
                for (int b = 0; b < TN; ++b) {
                    if (MODE == MODE_SUB) acc[a][b] = swap ? mf::msub(acc[a][b], bv[b], av[a]) : mf::msub(acc[a][b], av[a], bv[b]);
                    else acc[a][b] = mf::Mac<N, NET>::apply(acc[a][b], av[a], bv[b]);
                }
        }
        __syncthreads();
    }

    const mf::V<N> al = sv(alpha), be = sv(beta);
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) {
            const int64_t i = i0 + tx * TM + a, j = j0 + ty * TN + b;
            if (i >= row_hi || j >= n) continue;
            if (LOWER && i < j) continue;
            mf::V<N> r;
            if (MODE == MODE_SUB) r = acc[a][b];
            else if (beta_zero) r = mf::mul(al, acc[a][b]);
            else r = mf::add(mf::mul(al, acc[a][b]), mf::mul(be, ld<N>(C, i, j)));
            stv<N>(C, i, j, r);
        }
}

// ---------------------------------------------------------------- launcher
template <int N, int NET, int TM, int TN, int BK, int MODE, bool LOWER>
static hipError_t go(hipStream_t s, int64_t rows, int64_t n, int64_t K, const DV &A, const DV &B, const DV &C,
                     const S<N> &al, const S<N> &be, int bz, int64_t row_lo, const Status *st, int64_t lim, int swap) {
    constexpr int BM = 16 * TM, BN = 16 * TN;
    dim3 grid((unsigned)cdiv(rows, BM), (unsigned)cdiv(n, BN));
    hipLaunchKernelGGL((k_tile<N, NET, TM, TN, BK, MODE, LOWER>), grid, dim3(256), 0, s, n, K, A, B, C, al, be, bz, row_lo,
                       row_lo + rows, st, lim, swap);
    return hipGetLastError();
}

template <int N, int NET, int MODE, bool LOWER>
static hipError_t pick(hipStream_t s, int tile, int64_t rows, int64_t n, int64_t K, const DV &A, const DV &B, const DV &C,
                       const S<N> &al, const S<N> &be, int bz, int64_t row_lo, const Status *st, int64_t lim, int swap) {
    // tile 1: 32x32 tiles (2x2 per thread); tile 2: 32x64 tiles (2x4 per
    // thread, eight independent MAC chains per thread); tile 0: 16x16 tiles
    // (1x1 per thread) for small outputs where larger tiles would leave most
    // CUs idle.
    // tile 3: 64x64 tiles (4x4 per thread) for Float64x1/x2, whose short MAC
    // networks need sixteen chains per LDS read to keep the FP64 pipes busy.
    if constexpr (N <= 2) {
        if (tile == 3) return go<N, NET, 4, 4, 16, MODE, LOWER>(s, rows, n, K, A, B, C, al, be, bz, row_lo, st, lim, swap);
    }
    if (tile == 2 || tile == 3) return go<N, NET, 2, 4, 16, MODE, LOWER>(s, rows, n, K, A, B, C, al, be, bz, row_lo, st, lim, swap);
    if (tile == 1) return go<N, NET, 2, 2, 16, MODE, LOWER>(s, rows, n, K, A, B, C, al, be, bz, row_lo, st, lim, swap);
    return go<N, NET, 1, 1, 16, MODE, LOWER>(s, rows, n, K, A, B, C, al, be, bz, row_lo, st, lim, swap);
}

static int auto_tile(int N, int64_t rows, int64_t n, int64_t K, bool lower) {
    // Lower-triangle outputs (syrk!, gemmt!, the Cholesky / LDLT trailing
    // updates): 16x16 tiles. The triangle's tile count is rarely close to a
    // whole number of resident-workgroup rounds, and the small tiles shrink
    // the idle last round (x4 syrk n = 4096, k = 1024: 67 -> 60 ms; x4
    // Cholesky n = 4096, K = 16 updates: 115.6 ms with 32x32 tiles, 104.7 ms).
    if (lower) return 0;
    // Short reductions (the blocked factorizations' K = 16 trailing
    // updates): 16x16 tiles as well (x4 lu! n = 4096: 196.4 -> 193.9 ms).
    if (K <= 16) return 0;
    // 64x64 tiles for up to two limbs, 32x32 tiles otherwise, once they give
    // at least ~2 workgroups per CU (256 CUs).
    if (N <= 2 && cdiv(rows, 64) * cdiv(n, 64) >= 512) return 3;
    return (cdiv(rows, 32) * cdiv(n, 32) >= 512) ? 1 : 0;
}

int tile_geometry(int N, int64_t rows, int64_t n, int64_t K, bool lower, int requested, int *tile_rows, int *tile_cols) {
    int t = requested >= 0 ? requested : auto_tile(N, rows, n, K, lower);
    if (t < 0 || t > 3) t = 0;                 // pick(): values outside 0..3 run the 16x16 tile
    if (t == 3 && N > 2) t = 2;                // 64x64 exists for 1-2 limbs only
    static const int tm[4] = {1, 2, 2, 4}, tn[4] = {1, 2, 4, 4};
    *tile_rows = 16 * tm[t];
    *tile_cols = 16 * tn[t];
    return t;
}

template <int N>
hipError_t tile_launch(hipStream_t s, const TileArgs &t, const DV &A, const DV &B, const DV &C, const double *alpha,
                       const double *beta, const Status *st) {
    if (t.rows <= 0 || t.n <= 0) return hipSuccess;
    S<N> al, be;
    for (int l = 0; l < N; ++l) { al.x[l] = alpha ? alpha[l] : 0.0; be.x[l] = beta ? beta[l] : 0.0; }
    const int tile = t.tile >= 0 ? t.tile : auto_tile(N, t.rows, t.n, t.K, t.lower != 0);
    const int bz = t.beta_zero;
    if (t.mode == MODE_SUB) {
        if (t.lower) return pick<N, mf::NET_STD, MODE_SUB, true>(s, tile, t.rows, t.n, t.K, A, B, C, al, be, bz, t.row_lo, st, t.status_limit, t.swap_product);
        return pick<N, mf::NET_STD, MODE_SUB, false>(s, tile, t.rows, t.n, t.K, A, B, C, al, be, bz, t.row_lo, st, t.status_limit, t.swap_product);
    }
    if constexpr (N == 3) {
        if (t.net == mf::NET_FUSED3) {
            if (t.lower) return pick<3, mf::NET_FUSED3, MODE_AB, true>(s, tile, t.rows, t.n, t.K, A, B, C, al, be, bz, t.row_lo, st, t.status_limit, t.swap_product);
            return pick<3, mf::NET_FUSED3, MODE_AB, false>(s, tile, t.rows, t.n, t.K, A, B, C, al, be, bz, t.row_lo, st, t.status_limit, t.swap_product);
        }
    }
    if (t.lower) return pick<N, mf::NET_STD, MODE_AB, true>(s, tile, t.rows, t.n, t.K, A, B, C, al, be, bz, t.row_lo, st, t.status_limit, t.swap_product);
    return pick<N, mf::NET_STD, MODE_AB, false>(s, tile, t.rows, t.n, t.K, A, B, C, al, be, bz, t.row_lo, st, t.status_limit, t.swap_product);
}

// Narrow outputs (at most 8 columns): a workgroup of 256 threads takes 32
// rows x 8 columns, one output per thread, its operands read straight from
// memory (a(i, k) shared by the row's 8 threads, b(k, j) by the column's 32).
// Per element the operations and their order are k_tile's MODE_AB: the MAC
// chain over k ascending from zero, then alpha * acc + beta * C.
template <int N, bool LOWER>
__global__ __launch_bounds__(256) void k_cols(int64_t rows, int64_t ncols, int64_t K, DV A, DV B, DV C, S<N> alpha, S<N> beta,
                                              const Status *status, int64_t status_limit, const unsigned *wait_on, unsigned wait_value,
                                              unsigned *err) {
    const bool ok = wait_on == nullptr || wait_flag(wait_on, wait_value);
    if (status && status_skips(status, status_limit)) return;
    const int64_t i = (int64_t)blockIdx.x * 32 + (threadIdx.x & 31), j = threadIdx.x >> 5;
    if (i >= rows || j >= ncols || (LOWER && i < j)) return;
    if (!ok) {
        mf::V<N> bad;
#pragma unroll
        for (int l = 0; l < N; ++l) bad.x[l] = __builtin_nan("");
        stv<N>(C, i, j, bad);
        __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    mf::V<N> acc = mf::zero<N>();
    mf::V<N> a = K > 0 ? ld<N>(A, i, 0) : acc, b = K > 0 ? ld<N>(B, 0, j) : acc;
    for (int64_t k = 0; k < K; ++k) {
        const mf::V<N> ak = a, bk = b;
        if (k + 1 < K) {
            a = ld<N>(A, i, k + 1);
            b = ld<N>(B, k + 1, j);
        }
        acc = mf::Mac<N, mf::NET_STD>::apply(acc, ak, bk);
    }
    stv<N>(C, i, j, mf::add(mf::mul(sv(alpha), acc), mf::mul(sv(beta), ld<N>(C, i, j))));
}

hipError_t cols_launch_n(int N, hipStream_t s, int64_t rows, int64_t ncols, int64_t K, bool lower, const DV &A, const DV &B, const DV &C,
                         const double *alpha, const double *beta, const Status *st, int64_t status_limit, const unsigned *wait_on,
                         unsigned wait_value, unsigned *err) {
    if (wait_on && !err) return hipErrorInvalidValue;
    if (rows <= 0 || ncols <= 0) return hipSuccess;
    if (ncols > 8) return hipErrorInvalidValue;
    auto launch = [&](auto ntag) -> hipError_t {
        constexpr int NN = decltype(ntag)::value;
        S<NN> al, be;
        for (int l = 0; l < NN; ++l) { al.x[l] = alpha[l]; be.x[l] = beta[l]; }
        const dim3 grid((unsigned)cdiv(rows, 32));
        if (lower)
            hipLaunchKernelGGL((k_cols<NN, true>), grid, dim3(256), 0, s, rows, ncols, K, A, B, C, al, be, st, status_limit, wait_on, wait_value,
                               err);
        else
            hipLaunchKernelGGL((k_cols<NN, false>), grid, dim3(256), 0, s, rows, ncols, K, A, B, C, al, be, st, status_limit, wait_on,
                               wait_value, err);
        return hipGetLastError();
    };
    switch (N) {
        case 1: return launch(std::integral_constant<int, 1>());
        case 2: return launch(std::integral_constant<int, 2>());
        case 3: return launch(std::integral_constant<int, 3>());
        case 4: return launch(std::integral_constant<int, 4>());
    }
    return hipErrorInvalidValue;
}

hipError_t tile_launch_n(int N, hipStream_t s, const TileArgs &t, const DV &A, const DV &B, const DV &C, const double *alpha,
                         const double *beta, const Status *st) {
    switch (N) {
        case 1: return tile_launch<1>(s, t, A, B, C, alpha, beta, st);
        case 2: return tile_launch<2>(s, t, A, B, C, alpha, beta, st);
        case 3: return tile_launch<3>(s, t, A, B, C, alpha, beta, st);
        case 4: return tile_launch<4>(s, t, A, B, C, alpha, beta, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace mfk
