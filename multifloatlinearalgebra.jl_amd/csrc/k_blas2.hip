// Level-2 and elementwise Float64xN kernels: scaling, copies, row swaps,
// GEMV, SYMV, the mixed-precision residual, finiteness checks and the norm
// reductions behind normwise_backward_error.
#include "mf_common.hpp"
#include "k_blas2.hpp"

namespace mfk {

// ----------------------------------------------------------- elementwise
__global__ void k_status_reset(Status *st) {
    if (threadIdx.x == 0 && blockIdx.x == 0) { st->info = 0; st->aux = 0; st->sync = 0; st->sync_base = 0; }
}

template <int N>
__global__ void k_scale(int64_t m, int64_t n, DV B, S<N> alpha, const Status *st) {
    if (st && st->info != 0) return;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * n) return;
    const int64_t i = e % m, j = e / m;
    stv<N>(B, i, j, mf::mul(ld<N>(B, i, j), sv(alpha)));  // B[i] *= alpha (trsm.jl:10-16)
}

template <int N>
__global__ void k_zero(int64_t m, int64_t n, DV B) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * n) return;
    stv<N>(B, e % m, e / m, mf::zero<N>());
}

template <int N>
__global__ void k_copy(int64_t m, int64_t n, DV dst, DV src) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * n) return;
    const int64_t i = e % m, j = e / m;
    stv<N>(dst, i, j, ld<N>(src, i, j));
}

// Interleaved (AoS, Julia Matrix{Float64xN}: N limbs per element, column-major)
// <-> limb-planar conversion.
template <int N>
__global__ void k_aos_to_planar(int64_t m, int64_t n, DV dst, const double *src, int64_t src_ld) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * n) return;
    const int64_t i = e % m, j = e / m;
    mf::V<N> v;
#pragma unroll
    for (int l = 0; l < N; ++l) v.x[l] = src[(j * src_ld + i) * N + l];
    stv<N>(dst, i, j, v);
}
template <int N>
__global__ void k_planar_to_aos(int64_t m, int64_t n, double *dst, int64_t dst_ld, DV src) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * n) return;
    const int64_t i = e % m, j = e / m;
    mf::V<N> v = ld<N>(src, i, j);
#pragma unroll
    for (int l = 0; l < N; ++l) dst[(j * dst_ld + i) * N + l] = v.x[l];
}

// Sets st->info = -1 if any element of the selected region is non-finite.
// region: 0 all, 1 lower triangle (i >= j), 2 upper triangle (i <= j).
template <int N>
__global__ void k_finite(int64_t m, int64_t n, DV A, int region, Status *st) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * n) return;
    const int64_t i = e % m, j = e / m;
    if (region == 1 && i < j) return;
    if (region == 2 && i > j) return;
    if (!mf::is_finite(ld<N>(A, i, j))) st->info = -1;
}

// Elementwise networks over planar vectors (limb stride = count); used to pin
// the device arithmetic directly against fixed bit patterns.
template <int N> __host__ __device__ int arith_one(int op, const mf::V<N> &a, const mf::V<N> &b, const mf::V<N> &c, mf::V<N> &r) {
    switch (op) {
        case 0: r = mf::add(a, b); return 0;
        case 1: r = mf::sub(a, b); return 0;
        case 2: r = mf::mul(a, b); return 0;
        case 3: r = mf::div(a, b); return 0;
        case 4: r = mf::sqrt(a); return 0;
        case 5: r = mf::mac(c, a, b); return 0;
        case 6:
            if constexpr (N == 3) { r = mf::mac_fused3(c, a, b); return 0; }
            return -1;
        case 7: r = mf::inv(a); return 0;
        case 8: r = mf::msub(c, a, b); return 0;
        case 9: r = mf::vabs(a); return 0;
        case 10: r = mf::vmax(a, b); return 0;
        case 11: r = mf::vmin(a, b); return 0;
        case 12: r = mf::from<N>(mf::lt(a, b) ? 1.0 : 0.0); return 0;
    }
    return -1;
}

template <int N>
__global__ void k_arith(int op, int64_t count, const double *x, const double *y, const double *z, double *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    mf::V<N> a, b, c, r;
#pragma unroll
    for (int l = 0; l < N; ++l) {
        a.x[l] = x[l * count + i];
        b.x[l] = y ? y[l * count + i] : 0.0;
        c.x[l] = z ? z[l * count + i] : 0.0;
    }
    if (arith_one<N>(op, a, b, c, r) != 0) r = mf::from<N>(__builtin_nan(""));
#pragma unroll
    for (int l = 0; l < N; ++l) out[l * count + i] = r.x[l];
}

// --------------------------------------------------------------- GEMV
// src/kernels/gemv.jl:1-93: one thread per output, ascending reduction. The
// reduction is one serial chain per output (MFLA's order), so the kernel is
// bound by the chain's latency, not by HBM: 64-thread workgroups spread the
// outputs over as many CUs as possible, and a ring of GV_PF prefetched
// (A, x) pairs keeps the loads GV_PF steps ahead of the chain.
constexpr int GV_PF = 8;

template <int N>
__global__ __launch_bounds__(64) void k_gemv(int trans, int64_t outs, int64_t red, DV A, DV x, DV y, S<N> alpha, S<N> beta,
                                             int beta_zero, const Status *st) {
    if (st && st->info != 0) return;
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= outs) return;
    const DV Ao = trans ? DV{A.p + o * A.cs, A.cs, A.rs, A.ls} : DV{A.p + o * A.rs, A.rs, A.cs, A.ls};   // Ao(0, k)
    mf::V<N> acc = mf::zero<N>();
    mf::V<N> ra[GV_PF], rx[GV_PF];
#pragma unroll
    for (int q = 0; q < GV_PF; ++q) {
        if (q < red) { ra[q] = ld<N>(Ao, 0, q); rx[q] = ld<N>(x, q, 0); }
    }
    int64_t k = 0;
    for (; k + GV_PF <= red; k += GV_PF) {
#pragma unroll
        for (int q = 0; q < GV_PF; ++q) {
            acc = mf::mac(acc, ra[q], rx[q]);
            const int64_t kn = k + GV_PF + q;
            if (kn < red) { ra[q] = ld<N>(Ao, 0, kn); rx[q] = ld<N>(x, kn, 0); }
        }
    }
#pragma unroll
    for (int q = 0; q < GV_PF; ++q)
        if (k + q < red) acc = mf::mac(acc, ra[q], rx[q]);
    mf::V<N> r = beta_zero ? mf::mul(sv(alpha), acc) : mf::add(mf::mul(sv(alpha), acc), mf::mul(sv(beta), ld<N>(y, o, 0)));
    stv<N>(y, o, 0, r);
}

// src/kernels/symv.jl:1-36: columns 1..row-1, diagonal, row+1..n, reading
// only the authoritative triangle. One thread per output row.
template <int N>
__global__ __launch_bounds__(256) void k_symv(int upper, int64_t n, DV A, DV X, DV Y, int64_t nrhs, S<N> alpha, S<N> beta,
                                              int beta_zero) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * nrhs) return;
    const int64_t row = e % n, c = e / n;
    mf::V<N> acc = mf::zero<N>();
    for (int64_t col = 0; col < row; ++col) acc = mf::mac(acc, upper ? ld<N>(A, col, row) : ld<N>(A, row, col), ld<N>(X, col, c));
    acc = mf::mac(acc, ld<N>(A, row, row), ld<N>(X, row, c));
    for (int64_t col = row + 1; col < n; ++col) acc = mf::mac(acc, upper ? ld<N>(A, row, col) : ld<N>(A, col, row), ld<N>(X, col, c));
    mf::V<N> r = beta_zero ? mf::mul(sv(alpha), acc) : mf::add(mf::mul(sv(alpha), acc), mf::mul(sv(beta), ld<N>(Y, row, c)));
    stv<N>(Y, row, c, r);
}

// ------------------------------------------------- mixed residual
// src/residual.jl:365-409: acc = R(b); acc -= R(A)*R(x) ascending, all
// operands promoted (zero-extended limbs) to the residual precision.
template <int NS, int NR>
__global__ __launch_bounds__(256) void k_residual_mixed(int uplo, int64_t m, int64_t n, int64_t nrhs, DV A, DV X, DV B, DV R) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * nrhs) return;
    const int64_t row = e % m, c = e / m;
    auto prom = [](const mf::V<NS> &s) { mf::V<NR> r = mf::zero<NR>(); for (int l = 0; l < NS; ++l) r.x[l] = s.x[l]; return r; };
    mf::V<NR> acc = prom(ld<NS>(B, row, c));
    for (int64_t k = 0; k < n; ++k) {
        mf::V<NS> a;
        if (uplo == 0) a = ld<NS>(A, row, k);
        else if (uplo == 1) a = k <= row ? ld<NS>(A, row, k) : ld<NS>(A, k, row);
        else a = k >= row ? ld<NS>(A, row, k) : ld<NS>(A, k, row);
        acc = mf::msub(acc, prom(a), prom(ld<NS>(X, k, c)));
    }
    stv<NR>(R, row, c, acc);
}

// X (NR limbs) = X + promote(D) (NS <= NR limbs): correction accumulation of
// iterative refinement in the higher precision.
template <int NS, int NR>
__global__ void k_add_mixed(int64_t m, int64_t n, DV X, DV D) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * n) return;
    const int64_t i = e % m, j = e / m;
    mf::V<NS> d = ld<NS>(D, i, j);
    mf::V<NR> p = mf::zero<NR>();
#pragma unroll
    for (int l = 0; l < NS; ++l) p.x[l] = d.x[l];
    stv<NR>(X, i, j, mf::add(ld<NR>(X, i, j), p));
}

// ----------------------------------------------------- norm reductions
// Row sums of |A| (ascending column, src/residual.jl:143-164) per row.
template <int N>
__global__ void k_row_abs_sum(int uplo, int64_t m, int64_t n, DV A, DV out) {
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= m) return;
    mf::V<N> s = mf::zero<N>();
    for (int64_t c = 0; c < n; ++c) {
        mf::V<N> el;
        if (uplo == 0) el = ld<N>(A, row, c);
        else if (uplo == 1) el = c <= row ? ld<N>(A, row, c) : ld<N>(A, c, row);
        else el = c >= row ? ld<N>(A, row, c) : ld<N>(A, c, row);
        s = mf::add(s, mf::vabs(el));
    }
    stv<N>(out, row, 0, s);
}

// Column-wise max |X| and finiteness: out(0, c) = max_i |X(i, c)| in ascending
// i (Julia max semantics), fin[c] &= all finite. One workgroup per column.
template <int N>
__global__ void k_col_absmax(int64_t m, DV X, DV out, int *fin) {
    const int64_t c = blockIdx.x;
    __shared__ double red[256 * N];
    __shared__ int redf[256];
    mf::V<N> mx = mf::zero<N>();
    int f = 1;
    for (int64_t i = threadIdx.x; i < m; i += blockDim.x) {
        mf::V<N> v = ld<N>(X, i, c);
        if (!mf::is_finite(v)) f = 0;
        mx = mf::vmax(mx, mf::vabs(v));
    }
    for (int l = 0; l < N; ++l) red[threadIdx.x * N + l] = mx.x[l];
    redf[threadIdx.x] = f;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            mf::V<N> a, b;
            for (int l = 0; l < N; ++l) { a.x[l] = red[threadIdx.x * N + l]; b.x[l] = red[(threadIdx.x + s) * N + l]; }
            a = mf::vmax(a, b);
            for (int l = 0; l < N; ++l) red[threadIdx.x * N + l] = a.x[l];
            redf[threadIdx.x] = redf[threadIdx.x] & redf[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        mf::V<N> r;
        for (int l = 0; l < N; ++l) r.x[l] = red[l];
        stv<N>(out, c, 0, r);
        if (fin) fin[c] = redf[0];
    }
}

// ----------------------------------------------------------- launchers
static inline unsigned nb256(int64_t work) { return (unsigned)cdiv(work > 0 ? work : 1, 256); }

#define MF_DISPATCH(N, ...)                                        \
    switch (N) {                                                   \
        case 1: { constexpr int NN = 1; __VA_ARGS__; } break;      \
        case 2: { constexpr int NN = 2; __VA_ARGS__; } break;      \
        case 3: { constexpr int NN = 3; __VA_ARGS__; } break;      \
        case 4: { constexpr int NN = 4; __VA_ARGS__; } break;      \
        default: return hipErrorInvalidValue;                      \
    }

template <int N> static S<N> mk(const double *v) { S<N> s; for (int l = 0; l < N; ++l) s.x[l] = v ? v[l] : 0.0; return s; }

hipError_t arith(int N, hipStream_t s, int op, int64_t count, const double *x, const double *y, const double *z, double *out) {
    if (count == 0) return hipSuccess;
    MF_DISPATCH(N, hipLaunchKernelGGL(k_arith<NN>, dim3(nb256(count)), dim3(256), 0, s, op, count, x, y, z, out));
    return hipGetLastError();
}

int host_arith(int N, int op, int64_t count, const double *x, const double *y, const double *z, double *out) {
    MF_DISPATCH(N, {
        for (int64_t i = 0; i < count; ++i) {
            mf::V<NN> a, b, c, r;
            for (int l = 0; l < NN; ++l) {
                a.x[l] = x[l * count + i];
                b.x[l] = y ? y[l * count + i] : 0.0;
                c.x[l] = z ? z[l * count + i] : 0.0;
            }
            if (arith_one<NN>(op, a, b, c, r) != 0) return -1;
            for (int l = 0; l < NN; ++l) out[l * count + i] = r.x[l];
        }
    });
    return 0;
}

hipError_t status_reset(hipStream_t s, Status *st) {
    hipLaunchKernelGGL(k_status_reset, dim3(1), dim3(64), 0, s, st);
    return hipGetLastError();
}

hipError_t scale(int N, hipStream_t s, int64_t m, int64_t n, const DV &B, const double *alpha, const Status *st) {
    if (m * n == 0) return hipSuccess;
    MF_DISPATCH(N, hipLaunchKernelGGL(k_scale<NN>, dim3(nb256(m * n)), dim3(256), 0, s, m, n, B, mk<NN>(alpha), st));
    return hipGetLastError();
}

hipError_t zero_fill(int N, hipStream_t s, int64_t m, int64_t n, const DV &B) {
    if (m * n == 0) return hipSuccess;
    MF_DISPATCH(N, hipLaunchKernelGGL(k_zero<NN>, dim3(nb256(m * n)), dim3(256), 0, s, m, n, B));
    return hipGetLastError();
}

hipError_t copy(int N, hipStream_t s, int64_t m, int64_t n, const DV &dst, const DV &src) {
    if (m * n == 0) return hipSuccess;
    MF_DISPATCH(N, hipLaunchKernelGGL(k_copy<NN>, dim3(nb256(m * n)), dim3(256), 0, s, m, n, dst, src));
    return hipGetLastError();
}

hipError_t aos_to_planar(int N, hipStream_t s, int64_t m, int64_t n, const DV &dst, const double *src, int64_t src_ld) {
    if (m * n == 0) return hipSuccess;
    MF_DISPATCH(N, hipLaunchKernelGGL(k_aos_to_planar<NN>, dim3(nb256(m * n)), dim3(256), 0, s, m, n, dst, src, src_ld));
    return hipGetLastError();
}

hipError_t planar_to_aos(int N, hipStream_t s, int64_t m, int64_t n, double *dst, int64_t dst_ld, const DV &src) {
    if (m * n == 0) return hipSuccess;
    MF_DISPATCH(N, hipLaunchKernelGGL(k_planar_to_aos<NN>, dim3(nb256(m * n)), dim3(256), 0, s, m, n, dst, dst_ld, src));
    return hipGetLastError();
}

hipError_t finite_check(int N, hipStream_t s, int64_t m, int64_t n, const DV &A, int region, Status *st) {
    if (m * n == 0) return hipSuccess;
    MF_DISPATCH(N, hipLaunchKernelGGL(k_finite<NN>, dim3(nb256(m * n)), dim3(256), 0, s, m, n, A, region, st));
    return hipGetLastError();
}

hipError_t gemv(int N, hipStream_t s, int trans, int64_t m, int64_t n, const DV &A, const DV &x, const DV &y, const double *alpha,
                const double *beta, int beta_zero, const Status *st) {
    const int64_t outs = trans ? n : m, red = trans ? m : n;
    if (outs == 0) return hipSuccess;
    MF_DISPATCH(N, hipLaunchKernelGGL(k_gemv<NN>, dim3((unsigned)((outs + 63) / 64)), dim3(64), 0, s, trans, outs, red, A, x, y, mk<NN>(alpha),
                                      mk<NN>(beta), beta_zero, st));
    return hipGetLastError();
}

hipError_t symv(int N, hipStream_t s, int upper, int64_t n, int64_t nrhs, const DV &A, const DV &X, const DV &Y, const double *alpha,
                const double *beta, int beta_zero) {
    if (n * nrhs == 0) return hipSuccess;
    MF_DISPATCH(N, hipLaunchKernelGGL(k_symv<NN>, dim3(nb256(n * nrhs)), dim3(256), 0, s, upper, n, A, X, Y, nrhs, mk<NN>(alpha),
                                      mk<NN>(beta), beta_zero));
    return hipGetLastError();
}

hipError_t residual_mixed(int NS, int NR, hipStream_t s, int uplo, int64_t m, int64_t n, int64_t nrhs, const DV &A, const DV &X,
                          const DV &B, const DV &R) {
    if (m * nrhs == 0) return hipSuccess;
    const dim3 g(nb256(m * nrhs));
#define RM(a, b) hipLaunchKernelGGL((k_residual_mixed<a, b>), g, dim3(256), 0, s, uplo, m, n, nrhs, A, X, B, R)
    if (NS == 2 && NR == 3) RM(2, 3);
    else if (NS == 2 && NR == 4) RM(2, 4);
    else if (NS == 3 && NR == 4) RM(3, 4);
    else if (NS == 1 && NR == 2) RM(1, 2);
    else if (NS == 1 && NR == 3) RM(1, 3);
    else if (NS == 1 && NR == 4) RM(1, 4);
    else return hipErrorInvalidValue;
#undef RM
    return hipGetLastError();
}

hipError_t add_mixed(int NS, int NR, hipStream_t s, int64_t m, int64_t n, const DV &X, const DV &D) {
    if (m * n == 0) return hipSuccess;
    const dim3 g(nb256(m * n));
#define AM(a, b) hipLaunchKernelGGL((k_add_mixed<a, b>), g, dim3(256), 0, s, m, n, X, D)
    if (NS == 1 && NR == 2) AM(1, 2);
    else if (NS == 1 && NR == 3) AM(1, 3);
    else if (NS == 1 && NR == 4) AM(1, 4);
    else if (NS == 2 && NR == 3) AM(2, 3);
    else if (NS == 2 && NR == 4) AM(2, 4);
    else if (NS == 3 && NR == 4) AM(3, 4);
    else if (NS == NR && NS == 1) AM(1, 1);
    else if (NS == NR && NS == 2) AM(2, 2);
    else if (NS == NR && NS == 3) AM(3, 3);
    else if (NS == NR && NS == 4) AM(4, 4);
    else return hipErrorInvalidValue;
#undef AM
    return hipGetLastError();
}

hipError_t row_abs_sum(int N, hipStream_t s, int uplo, int64_t m, int64_t n, const DV &A, const DV &out) {
    if (m == 0) return hipSuccess;
    MF_DISPATCH(N, hipLaunchKernelGGL(k_row_abs_sum<NN>, dim3(nb256(m)), dim3(256), 0, s, uplo, m, n, A, out));
    return hipGetLastError();
}

hipError_t col_absmax(int N, hipStream_t s, int64_t m, int64_t ncols, const DV &X, const DV &out, int *fin) {
    if (ncols == 0) return hipSuccess;
    MF_DISPATCH(N, hipLaunchKernelGGL(k_col_absmax<NN>, dim3((unsigned)ncols), dim3(256), 0, s, m, X, out, fin));
    return hipGetLastError();
}

}  // namespace mfk
