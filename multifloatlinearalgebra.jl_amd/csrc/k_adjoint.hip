// Element kernels of the hand-written adjoints (solve, Cholesky, logdet).
// The matrix products and triangular solves of the adjoints reuse the
// Float64xN GEMM / TRSM kernels; what is here is the glue they need:
//
//  * k_symmetrize: out(i, j) = out(j, i) = alpha * (G(i, j) + G(j, i)) for
//    i >= j, one thread per pair, so the result is exactly symmetric (the
//    network evaluates G(i, j) + G(j, i) once per pair). In place allowed:
//    a pair's two entries are read and written by the same thread only.
//  * k_tril: keeps the lower triangle (optionally with the diagonal scaled
//    by 1/2, the Phi map of the Cholesky adjoint) and zeroes the rest.
//  * k_diag_logprod: the product of a triangle's diagonal as m * 2^e with m
//    in [1, 2), accumulated with exponent extraction so it neither overflows
//    nor underflows; 256 strided partial products, then a fixed-order tree.
#include "mf_common.hpp"
#include "k_adjoint.hpp"

namespace mfk {

template <int N>
__global__ void k_symmetrize(int64_t n, DV G, DV out, S<N> alpha) {
    const int64_t i = (int64_t)blockIdx.x * 16 + (threadIdx.x & 15);
    const int64_t j = (int64_t)blockIdx.y * 16 + (threadIdx.x >> 4);
    if (i >= n || j > i) return;
    const mf::V<N> s = mf::mul(sv(alpha), mf::add(ld<N>(G, i, j), ld<N>(G, j, i)));
    stv<N>(out, i, j, s);
    stv<N>(out, j, i, s);
}

template <int N>
__global__ void k_tril(int64_t n, DV src, DV dst, int halve_diag) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * n) return;
    const int64_t i = e % n, j = e / n;
    mf::V<N> v = mf::zero<N>();
    if (i >= j) {
        v = ld<N>(src, i, j);
        if (i == j && halve_diag) {
#pragma unroll
            for (int l = 0; l < N; ++l) v.x[l] *= 0.5;
        }
    }
    stv<N>(dst, i, j, v);
}

template <int N>
__global__ void k_identity(int64_t n, DV dst) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * n) return;
    const int64_t i = e % n, j = e / n;
    stv<N>(dst, i, j, mf::from<N>(i == j ? 1.0 : 0.0));
}

template <int N> __device__ __forceinline__ void renorm_exp(mf::V<N> &m, int64_t &e) {
    if (m.x[0] == 0.0 || !__builtin_isfinite(m.x[0])) return;
    const int k = ilogb(m.x[0]);
#pragma unroll
    for (int l = 0; l < N; ++l) m.x[l] = ldexp(m.x[l], -k);
    e += k;
}

// out: N limbs of m, then e (as a double, exact for |e| < 2^53)
template <int N>
__global__ __launch_bounds__(256) void k_diag_logprod(int64_t n, DV A, double *out) {
    __shared__ double ms[256 * N];
    __shared__ int64_t es[256];
    const int t = threadIdx.x;
    mf::V<N> m = mf::from<N>(1.0);
    int64_t e = 0;
    for (int64_t i = t; i < n; i += 256) {
        mf::V<N> d = ld<N>(A, i, i);
        renorm_exp(d, e);
        m = mf::mul(m, d);
        renorm_exp(m, e);
    }
#pragma unroll
    for (int l = 0; l < N; ++l) ms[t * N + l] = m.x[l];
    es[t] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
            mf::V<N> a, b;
#pragma unroll
            for (int l = 0; l < N; ++l) { a.x[l] = ms[t * N + l]; b.x[l] = ms[(t + s) * N + l]; }
            int64_t ee = es[t] + es[t + s];
            mf::V<N> p = mf::mul(a, b);
            renorm_exp(p, ee);
#pragma unroll
            for (int l = 0; l < N; ++l) ms[t * N + l] = p.x[l];
            es[t] = ee;
        }
        __syncthreads();
    }
    if (t == 0) {
#pragma unroll
        for (int l = 0; l < N; ++l) out[l] = ms[l];
        out[N] = (double)es[0];
    }
}

#define AD_DISPATCH(N, ...)                                        \
    switch (N) {                                                   \
        case 1: { constexpr int NN = 1; __VA_ARGS__; } break;      \
        case 2: { constexpr int NN = 2; __VA_ARGS__; } break;      \
        case 3: { constexpr int NN = 3; __VA_ARGS__; } break;      \
        case 4: { constexpr int NN = 4; __VA_ARGS__; } break;      \
        default: return hipErrorInvalidValue;                      \
    }

template <int N> S<N> mks(const double *a) { S<N> s; for (int l = 0; l < N; ++l) s.x[l] = a[l]; return s; }

hipError_t symmetrize(int N, hipStream_t s, int64_t n, const DV &G, const DV &out, const double *alpha) {
    if (n == 0) return hipSuccess;
    const dim3 grid((unsigned)cdiv(n, 16), (unsigned)cdiv(n, 16));
    AD_DISPATCH(N, hipLaunchKernelGGL(k_symmetrize<NN>, grid, dim3(256), 0, s, n, G, out, mks<NN>(alpha)));
    return hipGetLastError();
}

hipError_t tril(int N, hipStream_t s, int64_t n, const DV &src, const DV &dst, int halve_diag) {
    if (n == 0) return hipSuccess;
    AD_DISPATCH(N, hipLaunchKernelGGL(k_tril<NN>, dim3((unsigned)cdiv(n * n, 256)), dim3(256), 0, s, n, src, dst, halve_diag));
    return hipGetLastError();
}

hipError_t identity(int N, hipStream_t s, int64_t n, const DV &dst) {
    if (n == 0) return hipSuccess;
    AD_DISPATCH(N, hipLaunchKernelGGL(k_identity<NN>, dim3((unsigned)cdiv(n * n, 256)), dim3(256), 0, s, n, dst));
    return hipGetLastError();
}

hipError_t diag_logprod(int N, hipStream_t s, int64_t n, const DV &A, double *out) {
    AD_DISPATCH(N, hipLaunchKernelGGL(k_diag_logprod<NN>, dim3(1), dim3(256), 0, s, n, A, out));
    return hipGetLastError();
}

}  // namespace mfk
