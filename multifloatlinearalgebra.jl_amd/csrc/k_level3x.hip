// trmm!, mfdot and the packed-storage moves of syrk_packed!.
//
//  * k_trmm: C(i, j) = alpha * acc with acc reduced in MFLA's order
//    (src/kernels/trmm.jl:1-73). Written as a triangular GEMM over a copy of
//    B: left C = op(A) * Bc with k in [0, i] (lower, diagonal last) or
//    [i, m) (upper, diagonal first); right C = Bc * op(A) with k in [j, n)
//    (lower, diagonal first) or [0, j] (upper, diagonal last). Each term is
//    acc + x*y with MFLA's operand order (coeff*B left, B*coeff right); a
//    unit diagonal contributes acc + B(i, j) instead. 32x32 output tiles,
//    k-panels of 16 staged in LDS; the k range of a tile is clipped to its
//    triangle and each element skips the terms outside its own range.
//  * k_dot: mfdot's single ascending chain over the column-major linear
//    index (src/kernels/dot.jl:8-22). The chain is serial by definition; the
//    workgroup stages chunks of x and y in LDS and one lane runs the chain.
//  * k_pack_lower / k_unpack_lower: dense lower triangle <-> packed
//    column-major lower storage, (r, c) -> c(2n - c + 1)/2 + (r - c).
#include "mf_common.hpp"
#include "k_level3x.hpp"

namespace mfk {

constexpr int TR_T = 32, TR_BK = 16;

template <int N>
__global__ __launch_bounds__(256) void k_trmm(int64_t m, int64_t n, int64_t K, DV X, DV Y, DV C, S<N> alpha, int left, int lower,
                                              int unit) {
    __shared__ double Xs[TR_BK * TR_T * N];
    __shared__ double Ys[TR_BK * TR_T * N];
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const int64_t i0 = (int64_t)blockIdx.x * TR_T, j0 = (int64_t)blockIdx.y * TR_T;
    // k range of the tile
    int64_t klo = 0, khi = K;
    if (left) {
        if (lower) khi = (i0 + TR_T < K) ? i0 + TR_T : K;
        else klo = i0;
    } else {
        if (lower) klo = j0;
        else khi = (j0 + TR_T < K) ? j0 + TR_T : K;
    }
    mf::V<N> acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = mf::zero<N>();

    for (int64_t k0 = klo; k0 < khi; k0 += TR_BK) {
        for (int e = tid; e < TR_T * TR_BK; e += 256) {
            const int ii = e % TR_T, kk = e / TR_T;
            const int64_t gi = i0 + ii, gk = k0 + kk;
            const bool ok = gi < m && gk < khi;
#pragma unroll
            for (int l = 0; l < N; ++l) Xs[(kk * TR_T + ii) * N + l] = ok ? X.p[gi * X.rs + gk * X.cs + l * X.ls] : 0.0;
        }
        for (int e = tid; e < TR_T * TR_BK; e += 256) {
            const int jj = e % TR_T, kk = e / TR_T;
            const int64_t gj = j0 + jj, gk = k0 + kk;
            const bool ok = gj < n && gk < khi;
#pragma unroll
            for (int l = 0; l < N; ++l) Ys[(kk * TR_T + jj) * N + l] = ok ? Y.p[gk * Y.rs + gj * Y.cs + l * Y.ls] : 0.0;
        }
        __syncthreads();
        const int kend = (khi - k0) < TR_BK ? (int)(khi - k0) : TR_BK;
        for (int kk = 0; kk < kend; ++kk) {
            const int64_t k = k0 + kk;
            mf::V<N> xv[2], yv[2];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int l = 0; l < N; ++l) xv[a].x[l] = Xs[(kk * TR_T + tx * 2 + a) * N + l];
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int l = 0; l < N; ++l) yv[b].x[l] = Ys[(kk * TR_T + ty * 2 + b) * N + l];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    const int64_t i = i0 + tx * 2 + a, j = j0 + ty * 2 + b;
                    const int64_t d = left ? i : j;
                    const bool in = left ? (lower ? k <= i : k >= i) : (lower ? k >= j : k <= j);
                    if (!in) continue;
                    if (unit && k == d) acc[a][b] = mf::add(acc[a][b], left ? yv[b] : xv[a]);
                    else acc[a][b] = mf::add(acc[a][b], mf::mul(xv[a], yv[b]));
                }
        }
        __syncthreads();
    }
    const mf::V<N> al = sv(alpha);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int64_t i = i0 + tx * 2 + a, j = j0 + ty * 2 + b;
            if (i < m && j < n) stv<N>(C, i, j, mf::mul(al, acc[a][b]));
        }
}

constexpr int DOT_CHUNK = 1024;

template <int N>
__global__ __launch_bounds__(256) void k_dot(int64_t rows, int64_t cols, DV x, DV y, double *out) {
    __shared__ double xs[DOT_CHUNK * N];
    __shared__ double ys[DOT_CHUNK * N];
    const int64_t total = rows * cols;
    mf::V<N> acc = mf::zero<N>();
    for (int64_t base = 0; base < total; base += DOT_CHUNK) {
        const int cnt = (total - base) < DOT_CHUNK ? (int)(total - base) : DOT_CHUNK;
        for (int e = threadIdx.x; e < cnt; e += blockDim.x) {
            const int64_t li = base + e, i = li % rows, j = li / rows;
#pragma unroll
            for (int l = 0; l < N; ++l) {
                xs[e * N + l] = x.p[i * x.rs + j * x.cs + l * x.ls];
                ys[e * N + l] = y.p[i * y.rs + j * y.cs + l * y.ls];
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int e = 0; e < cnt; ++e) {
                mf::V<N> a, b;
#pragma unroll
                for (int l = 0; l < N; ++l) { a.x[l] = xs[e * N + l]; b.x[l] = ys[e * N + l]; }
                acc = mf::add(acc, mf::mul(a, b));
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int l = 0; l < N; ++l) out[l] = acc.x[l];
    }
}

// packed index of (r, c), r >= c, in an n x n lower triangle
__device__ __forceinline__ int64_t packed_index(int64_t n, int64_t r, int64_t c) { return c * (2 * n - c + 1) / 2 + (r - c); }

template <int N>
__global__ void k_pack_lower(int64_t n, DV dense, double *packed, int64_t pls, int unpack) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * n) return;
    const int64_t r = e % n, c = e / n;
    if (r < c) return;
    const int64_t q = packed_index(n, r, c);
    double *d = dense.p + r * dense.rs + c * dense.cs;
#pragma unroll
    for (int l = 0; l < N; ++l) {
        if (unpack) d[l * dense.ls] = packed[l * pls + q];
        else packed[l * pls + q] = d[l * dense.ls];
    }
}

#define L3X_DISPATCH(N, ...)                                       \
    switch (N) {                                                   \
        case 1: { constexpr int NN = 1; __VA_ARGS__; } break;      \
        case 2: { constexpr int NN = 2; __VA_ARGS__; } break;      \
        case 3: { constexpr int NN = 3; __VA_ARGS__; } break;      \
        case 4: { constexpr int NN = 4; __VA_ARGS__; } break;      \
        default: return hipErrorInvalidValue;                      \
    }

template <int N> static S<N> mk_s(const double *a) { S<N> s; for (int l = 0; l < N; ++l) s.x[l] = a[l]; return s; }

hipError_t trmm(int N, hipStream_t s, int left, int lower, int unit, int64_t m, int64_t n, int64_t K, const DV &X, const DV &Y,
                const DV &C, const double *alpha) {
    if (m == 0 || n == 0) return hipSuccess;
    const dim3 grid((unsigned)cdiv(m, TR_T), (unsigned)cdiv(n, TR_T));
    L3X_DISPATCH(N, hipLaunchKernelGGL(k_trmm<NN>, grid, dim3(256), 0, s, m, n, K, X, Y, C, mk_s<NN>(alpha), left, lower, unit));
    return hipGetLastError();
}

hipError_t dot(int N, hipStream_t s, int64_t rows, int64_t cols, const DV &x, const DV &y, double *out) {
    L3X_DISPATCH(N, hipLaunchKernelGGL(k_dot<NN>, dim3(1), dim3(256), 0, s, rows, cols, x, y, out));
    return hipGetLastError();
}

hipError_t pack_lower(int N, hipStream_t s, int64_t n, const DV &dense, double *packed, int64_t pls, int unpack) {
    if (n == 0) return hipSuccess;
    L3X_DISPATCH(N, hipLaunchKernelGGL(k_pack_lower<NN>, dim3((unsigned)cdiv(n * n, 256)), dim3(256), 0, s, n, dense, packed, pls, unpack));
    return hipGetLastError();
}

}  // namespace mfk
