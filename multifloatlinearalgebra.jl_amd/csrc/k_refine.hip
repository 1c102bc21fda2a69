// Kernels of the FP64-factor + Float64xN iterative refinement driver.
//
//  * k_trinv_blocks: inverts the nb x nb diagonal blocks of a triangular
//    factor (systolic solve of each block against the identity, all blocks in
//    one launch), so a correction solve becomes n/nb short block GEMMs
//    instead of an n-step serial substitution chain.
//  * k_resid_update<NA, NR>: R(NR limbs) -= sum_k A(i,k) * D(k,c), with A
//    truncated to its leading NA limbs, D an FP64 correction, and the product
//    sum accumulated in NA-limb arithmetic. The refinement picks NA from the
//    measured size of the correction: a correction that is 2^-53j of the
//    solution only needs its product with A to ~u_N / 2^-53j relative
//    accuracy, so later (smaller) corrections use shorter networks.
#include "mf_common.hpp"
#include "k_refine.hpp"

namespace mfk {

// ------------------------------------------------------- block inverses
// Block b (rows/cols [b*nb, b*nb + nbk)) of the triangle in T is inverted into
// the same position of W (unit triangles get an explicit unit diagonal).
template <int N>
__global__ __launch_bounds__(256) void k_trinv_blocks(int64_t n, int nb, DV T, DV W, int lower, int unit) {
    __shared__ double xs[4 * N];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;   // row in block, RHS in group
    const int64_t b = blockIdx.y;
    const int64_t p0 = b * nb;
    const int nbk = (int)((n - p0) < nb ? (n - p0) : nb);
    const int c = (int)blockIdx.x * 4 + ty;                   // RHS = identity column in block
    const bool valid = tx < nbk && c < nbk;
    mf::V<N> v = mf::from<N>((valid && tx == c) ? 1.0 : 0.0);
    for (int s0 = 0; s0 < nbk; ++s0) {
        const int s = lower ? s0 : nbk - 1 - s0;
        if (tx == s && valid) {
            if (!unit) v = mf::div(v, ld<N>(T, p0 + s, p0 + s));
            stv<N>(W, p0 + s, p0 + c, v);
#pragma unroll
            for (int l = 0; l < N; ++l) xs[ty * N + l] = v.x[l];
        }
        __syncthreads();
        if (valid && (lower ? tx > s : tx < s)) {
            mf::V<N> xq;
#pragma unroll
            for (int l = 0; l < N; ++l) xq.x[l] = xs[ty * N + l];
            v = mf::msub(v, ld<N>(T, p0 + tx, p0 + s), xq);
        }
        __syncthreads();
    }
}

// --------------------------------------------- mixed-precision products
// a (NA limbs) * d (binary64) to NA-limb accuracy (mf::mul_dbl).
template <int NA> __device__ __forceinline__ mf::V<NA> mul_d(const mf::V<NA> &a, double d) { return mf::mul_dbl<NA>(a, d); }

// R(NR) -= A(top NA limbs) * D(FP64); 16x16 output tile per 256 threads, A and
// D k-panels staged in LDS, ascending k.
// uplo 1 / 2: A is symmetric and only its lower / upper triangle is read.
template <int NA, int NR>
__global__ __launch_bounds__(256) void k_resid_update(int64_t m, int64_t nrhs, int64_t K, DV A, DV D, DV R, int uplo) {
    constexpr int BK = 32;
    __shared__ double As[BK * 16 * NA];
    __shared__ double Ds[BK * 16];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int64_t i = (int64_t)blockIdx.x * 16 + tx, c = (int64_t)blockIdx.y * 16 + ty;
    mf::V<NA> acc = mf::zero<NA>();
    for (int64_t k0 = 0; k0 < K; k0 += BK) {
        for (int e = threadIdx.x; e < 16 * BK; e += 256) {
            const int ii = e % 16, kk = e / 16;
            int64_t gi = (int64_t)blockIdx.x * 16 + ii, gk = k0 + kk;
            const bool ok = gi < m && gk < K;
            if ((uplo == 1 && gi < gk) || (uplo == 2 && gi > gk)) { const int64_t q = gi; gi = gk; gk = q; }
#pragma unroll
            for (int l = 0; l < NA; ++l) As[(kk * 16 + ii) * NA + l] = ok ? A.p[gi * A.rs + gk * A.cs + l * A.ls] : 0.0;
        }
        for (int e = threadIdx.x; e < 16 * BK; e += 256) {
            const int kk = e % BK, jj = e / BK;
            const int64_t gj = (int64_t)blockIdx.y * 16 + jj, gk = k0 + kk;
            Ds[kk * 16 + jj] = (gj < nrhs && gk < K) ? D.p[gk * D.rs + gj * D.cs] : 0.0;
        }
        __syncthreads();
        const int kend = (K - k0) < BK ? (int)(K - k0) : BK;
        for (int kk = 0; kk < kend; ++kk) {
            mf::V<NA> a;
#pragma unroll
            for (int l = 0; l < NA; ++l) a.x[l] = As[(kk * 16 + tx) * NA + l];
            acc = mf::add(acc, mul_d<NA>(a, Ds[kk * 16 + ty]));
        }
        __syncthreads();
    }
    if (i < m && c < nrhs) {
        mf::V<NR> p = mf::zero<NR>();
#pragma unroll
        for (int l = 0; l < NA; ++l) p.x[l] = acc.x[l];
        stv<NR>(R, i, c, mf::sub(ld<NR>(R, i, c), p));
    }
}

// Column max |X| (limb 0 only, for the precision choice) of an m x ncols matrix.
__global__ void k_col_absmax_hi(int64_t m, DV X, double *out) {
    __shared__ double red[256];
    const int64_t c = blockIdx.x;
    double mx = 0.0;
    for (int64_t i = threadIdx.x; i < m; i += blockDim.x) mx = fmax(mx, fabs(X.p[i * X.rs + c * X.cs]));
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = red[0];
}

// ---------------------------------------------------------- launchers
hipError_t trinv_blocks(int N, hipStream_t s, int64_t n, int nb, const DV &T, const DV &W, int lower, int unit) {
    if (n == 0) return hipSuccess;
    if (nb > 64) return hipErrorInvalidValue;
    const dim3 grid((unsigned)cdiv(nb, 4), (unsigned)cdiv(n, nb));
    switch (N) {
        case 1: hipLaunchKernelGGL(k_trinv_blocks<1>, grid, dim3(256), 0, s, n, nb, T, W, lower, unit); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t resid_update(int NA, int NR, hipStream_t s, int uplo, int64_t m, int64_t nrhs, int64_t K, const DV &A, const DV &D, const DV &R) {
    if (m == 0 || nrhs == 0) return hipSuccess;
    const dim3 grid((unsigned)cdiv(m, 16), (unsigned)cdiv(nrhs, 16));
#define RU(a, b) hipLaunchKernelGGL((k_resid_update<a, b>), grid, dim3(256), 0, s, m, nrhs, K, A, D, R, uplo)
    if (NR == 4) {
        if (NA == 4) RU(4, 4); else if (NA == 3) RU(3, 4); else if (NA == 2) RU(2, 4); else RU(1, 4);
    } else if (NR == 3) {
        if (NA == 3) RU(3, 3); else if (NA == 2) RU(2, 3); else RU(1, 3);
    } else if (NR == 2) {
        if (NA == 2) RU(2, 2); else RU(1, 2);
    } else {
        return hipErrorInvalidValue;
    }
#undef RU
    return hipGetLastError();
}

hipError_t col_absmax_hi(hipStream_t s, int64_t m, int64_t ncols, const DV &X, double *out) {
    if (ncols == 0) return hipSuccess;
    hipLaunchKernelGGL(k_col_absmax_hi, dim3((unsigned)ncols), dim3(256), 0, s, m, X, out);
    return hipGetLastError();
}

}  // namespace mfk
