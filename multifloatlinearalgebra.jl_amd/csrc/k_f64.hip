// Binary64 kernels of the refinement's low-precision factor and correction
// solve. These run in plain FP64 (no multi-double networks, no MFLA order
// constraints), so the GEMM-shaped work goes to the CDNA4 FP64 matrix cores:
//
//  * k_dgemm: C = alpha * op(A) * op(B) + beta * C on strided views, 64x64
//    tile per 256-thread workgroup (four waves, each 32x32 = 2x2 MFMA
//    16x16x4 f64 blocks), A/B k-panels of 16 staged in LDS. Optional
//    triangular structure of A / B (the k range of each tile is clipped and
//    the zero triangle is never read), a lower-triangular output mask, a
//    batch index (grid z) with element strides, and a split-k mode that
//    writes partial sums for k_dgemm_reduce (fixed summation order, so the
//    result is deterministic).
//  * k_dpotrf_diag: Cholesky of one nb x nb (nb <= 64) diagonal block in LDS,
//    plus the inverse of the resulting triangle (needed by the panel update
//    L21 = A21 * inv(L11)^T and by the correction solve).
//  * small helpers: copy of a binary64 plane with optional mirroring, row
//    sums of |A|, row gathers by a permutation.
#include "mf_common.hpp"
#include "k_f64.hpp"

namespace mfk {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int DG_T = 64, DG_BK = 16, DG_PAD = 1;

// v_mfma_f64_16x16x4_f64 operand layout (lane l of the wave), as measured
// on gfx950 with integer-valued operands:
//   first operand:  P[l % 16][l / 16]   (16 x 4)
//   second operand: Q[l / 16][l % 16]   (4 x 16)
//   D item v:       D[l / 16 + 4 * v][l % 16]   (D = P Q)
// The kernel forms the transposed tile D = B^T A^T (P = B^T, Q = A^T), so
// lane l holds C rows i = l % 16: the 16 lanes of a row group store 16
// consecutive elements of a column-major C column.
__global__ __launch_bounds__(256) void k_dgemm(DGemmArgs g, int tiles_m) {
    __shared__ double As[DG_BK][DG_T + DG_PAD];
    __shared__ double Bs[DG_BK][DG_T + DG_PAD];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wy = w & 1, wx = w >> 1;
    const int64_t tm = blockIdx.x % tiles_m, tn = blockIdx.x / tiles_m;
    const int64_t i0 = tm * DG_T, j0 = tn * DG_T;
    const int64_t z = blockIdx.z;
    const double *Ap = g.A.p + z * g.za, *Bp = g.B.p + z * g.zb;
    double *Cp = g.C.p + z * g.zc;
    if (g.c_lower && j0 > i0 + DG_T - 1) return;  // tile entirely above the diagonal

    // k range of this tile (zero triangles of A / B contribute nothing)
    int64_t klo = 0, khi = g.k;
    if (g.a_tri == TRI_LOWER) khi = khi < i0 + DG_T ? khi : i0 + DG_T;
    if (g.a_tri == TRI_UPPER) klo = klo > i0 ? klo : i0;
    if (g.b_tri == TRI_LOWER) klo = klo > j0 ? klo : j0;
    if (g.b_tri == TRI_UPPER) khi = khi < j0 + DG_T ? khi : j0 + DG_T;
    // split-k: slice s of ksplit covers an equal share of [klo, khi)
    const int ks = blockIdx.y;
    if (g.ksplit > 1) {
        const int64_t span = khi > klo ? khi - klo : 0;
        const int64_t chunk = ((span + g.ksplit - 1) / g.ksplit + DG_BK - 1) / DG_BK * DG_BK;
        const int64_t a = klo + ks * chunk, b = a + chunk;
        klo = a;
        khi = b < khi ? b : khi;
    }

    d4 acc[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = d4{0.0, 0.0, 0.0, 0.0};

    const bool a_rows_fast = g.A.rs == 1, b_cols_fast = g.B.cs == 1;
    for (int64_t k0 = klo; k0 < khi; k0 += DG_BK) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int ii, kk;
            if (a_rows_fast) { ii = t & 63; kk = (t >> 6) + 4 * q; }
            else { kk = t & 15; ii = (t >> 4) + 16 * q; }
            const int64_t gi = i0 + ii, gk = k0 + kk;
            bool ok = gi < g.m && gk < khi;
            if (g.a_tri == TRI_LOWER) ok = ok && gk <= gi;
            if (g.a_tri == TRI_UPPER) ok = ok && gk >= gi;
            As[kk][ii] = ok ? Ap[gi * g.A.rs + gk * g.A.cs] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int jj, kk;
            if (b_cols_fast) { jj = t & 63; kk = (t >> 6) + 4 * q; }
            else { kk = t & 15; jj = (t >> 4) + 16 * q; }
            const int64_t gj = j0 + jj, gk = k0 + kk;
            bool ok = gj < g.n && gk < khi;
            if (g.b_tri == TRI_LOWER) ok = ok && gk >= gj;
            if (g.b_tri == TRI_UPPER) ok = ok && gk <= gj;
            Bs[kk][jj] = ok ? Bp[gk * g.B.rs + gj * g.B.cs] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int k4 = 0; k4 < DG_BK / 4; ++k4) {
            const int kr = k4 * 4 + (lane >> 4);
            double a[2], b[2];
#pragma unroll
            for (int bi = 0; bi < 2; ++bi) a[bi] = As[kr][wy * 32 + bi * 16 + (lane & 15)];
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) b[bj] = Bs[kr][wx * 32 + bj * 16 + (lane & 15)];
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[bj], a[bi], acc[bi][bj], 0, 0, 0);
        }
        __syncthreads();
    }

#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t i = i0 + wy * 32 + bi * 16 + (lane & 15);
                const int64_t j = j0 + wx * 32 + bj * 16 + (lane >> 4) + 4 * v;
                if (i >= g.m || j >= g.n) continue;
                if (g.c_lower && i < j) continue;
                const double r = acc[bi][bj][v];
                if (g.ksplit > 1) {
                    g.P[((int64_t)ks * g.n + j) * g.m + i] = r;
                } else {
                    double *c = Cp + i * g.C.rs + j * g.C.cs;
                    *c = g.beta_zero ? g.alpha * r : g.alpha * r + g.beta * *c;
                }
            }
}

// C = alpha * sum_s P[s] + beta * C, slices summed in ascending s.
__global__ void k_dgemm_reduce(DGemmArgs g) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.m * g.n) return;
    const int64_t i = e % g.m, j = e / g.m;
    if (g.c_lower && i < j) return;
    double s = 0.0;
    for (int q = 0; q < g.ksplit; ++q) s += g.P[((int64_t)q * g.n + j) * g.m + i];
    double *c = g.C.p + i * g.C.rs + j * g.C.cs;
    *c = g.beta_zero ? g.alpha * s : g.alpha * s + g.beta * *c;
}

// ------------------------------------------------ diagonal block Cholesky
// Factor the nb x nb block at (j0, j0) of F (lower triangle read), write L
// back to F's lower triangle and inv(L) to Inv's lower triangle at (j0, j0).
// The block is padded to 64 x 64 with an identity. Factorization: right-
// looking with deferred scaling (a_rq -= a_rc a_qc / d_c, one barrier per
// column), then L = column c / sqrt(d_c). Inverse: recursive doubling over
// block sizes 1, 2, 4, .., 32, each level two small triangular products
// (W = L21 X11 kept in the strict upper triangle, X21 = -X22 W). A
// non-positive or non-finite pivot at block column c records info =
// j0 + c + 1 and stops the factorization.
constexpr int DP_LD = 65, DP_THREADS = 1024;

__global__ __launch_bounds__(DP_THREADS) void k_dpotrf_diag(DV F, DV Inv, int64_t j0, int nb, Status *st) {
    if (st->info != 0) return;
    __shared__ double L[64 * DP_LD];   // column-major: L[q * DP_LD + r]
    __shared__ double X[64 * DP_LD];   // row-major:    X[r * DP_LD + c]
    __shared__ double sd[64];
    const int t = threadIdx.x;
    for (int e = t; e < 64 * 64; e += DP_THREADS) {
        const int r = e & 63, q = e >> 6;
        double v = (r == q) ? 1.0 : 0.0;
        if (r < nb && q < nb && r >= q) v = F.p[(j0 + r) * F.rs + (j0 + q) * F.cs];
        L[q * DP_LD + r] = v;
    }
    __syncthreads();
    for (int c = 0; c < nb; ++c) {
        const double d = L[c * DP_LD + c];
        if (!(d > 0.0) || !__builtin_isfinite(d)) {      // uniform: every thread reads the same pivot
            if (t == 0) st->info = j0 + c + 1;
            return;
        }
        const double inv_d = 1.0 / d;
        const int w = nb - c - 1;
        for (int e = t; e < w * w; e += DP_THREADS) {
            const int r = c + 1 + e % w, q = c + 1 + e / w;
            if (r >= q) L[q * DP_LD + r] -= L[c * DP_LD + r] * L[c * DP_LD + q] * inv_d;
        }
        __syncthreads();
    }
    if (t < 64) sd[t] = __builtin_sqrt(L[t * DP_LD + t]);
    __syncthreads();
    for (int e = t; e < 64 * 64; e += DP_THREADS) {
        const int r = e & 63, q = e >> 6;
        if (r > q) L[q * DP_LD + r] /= sd[q];
    }
    __syncthreads();
    if (t < 64) {
        L[t * DP_LD + t] = sd[t];
        X[t * DP_LD + t] = 1.0 / sd[t];
    }
    __syncthreads();
    for (int e = t; e < 64 * 64; e += DP_THREADS) {
        const int r = e & 63, q = e >> 6;
        if (r < nb && q < nb && r >= q) F.p[(j0 + r) * F.rs + (j0 + q) * F.cs] = L[q * DP_LD + r];
    }
    for (int s = 1; s < 64; s *= 2) {
        const int pairs = 64 / (2 * s);
        for (int e = t; e < pairs * s * s; e += DP_THREADS) {       // W = L21 X11 (X11 lower: k >= j)
            const int p = e / (s * s), i = (e / s) % s, j = e % s, b = p * 2 * s;
            double acc = 0.0;
            for (int k = j; k < s; ++k) acc += L[(b + k) * DP_LD + b + s + i] * X[(b + k) * DP_LD + b + j];
            X[(b + j) * DP_LD + b + s + i] = acc;
        }
        __syncthreads();
        for (int e = t; e < pairs * s * s; e += DP_THREADS) {       // X21 = -X22 W (X22 lower: k <= i)
            const int p = e / (s * s), i = (e / s) % s, j = e % s, b = p * 2 * s;
            double acc = 0.0;
            for (int k = 0; k <= i; ++k) acc += X[(b + s + i) * DP_LD + b + s + k] * X[(b + j) * DP_LD + b + s + k];
            X[(b + s + i) * DP_LD + b + j] = -acc;
        }
        __syncthreads();
    }
    for (int e = t; e < 64 * 64; e += DP_THREADS) {
        const int r = e & 63, q = e >> 6;
        if (r < nb && q < nb && r >= q) Inv.p[(j0 + r) * Inv.rs + (j0 + q) * Inv.cs] = X[r * DP_LD + q];
    }
}

// ------------------------------------------------------------ helpers
// dst(i, j) = src(i, j) for the requested region; mirror = 1 fills the upper
// triangle from the lower one (dst(i, j) = src(j, i) for i < j), mirror = 2
// the lower from the upper.
__global__ void k_dcopy(int64_t m, int64_t n, DV dst, DV src, int mirror) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * n) return;
    const int64_t i = e % m, j = e / m;
    int64_t si = i, sj = j;
    if ((mirror == 1 && i < j) || (mirror == 2 && i > j)) { si = j; sj = i; }
    dst.p[i * dst.rs + j * dst.cs] = src.p[si * src.rs + sj * src.cs];
}

// part[c * m + i] = sum over column chunk c of |A(i, j)| (binary64 plane;
// uplo 1 / 2 reads a stored triangle symmetrically); k_drow_abs_finish adds
// the chunks of each row in ascending order.
constexpr int ROW_CHUNKS = 32;

__global__ void k_drow_abs_sum(int64_t m, int64_t n, DV A, int uplo, double *part) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int64_t c = blockIdx.y, per = (n + ROW_CHUNKS - 1) / ROW_CHUNKS;
    const int64_t jlo = c * per, jhi = (jlo + per < n) ? jlo + per : n;
    double s = 0.0;
    for (int64_t j = jlo; j < jhi; ++j) {
        int64_t a = i, b = j;
        if ((uplo == 1 && a < b) || (uplo == 2 && a > b)) { a = j; b = i; }
        s += fabs(A.p[a * A.rs + b * A.cs]);
    }
    part[c * m + i] = s;
}

__global__ void k_drow_abs_finish(int64_t m, const double *part, double *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double s = 0.0;
    for (int c = 0; c < ROW_CHUNKS; ++c) s += part[c * m + i];
    out[i] = s;
}

// dst(i, c) = src(perm[i], c)
__global__ void k_dgather_rows(int64_t m, int64_t ncols, DV dst, DV src, const int64_t *perm) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * ncols) return;
    const int64_t i = e % m, c = e / m;
    dst.p[i * dst.rs + c * dst.cs] = src.p[perm[i] * src.rs + c * src.cs];
}

// ------------------------------------------------- binary64 LU (leaves)
// k_dlu_leaf: partial-pivoting LU of the w (<= W) columns [c0, c0 + w) over
// rows [c0, m) by one 1024-thread workgroup with the panel in registers:
// thread t holds rows c0 + t + 1024 i (i < RPT). Per column: workgroup argmax
// of |a| (first maximum on ties), exchange of the pivot row through LDS,
// scaling by the reciprocal pivot and a rank-1 update of the leaf's later
// columns. ipiv gets 1-based global row indices. A zero (or NaN-only) pivot
// column records info = column + 1 if no failure is recorded yet and is
// skipped, as dgetf2 skips it.
constexpr int LU_THREADS = 1024, LU_WAVES = LU_THREADS / 64;

__device__ __forceinline__ void lu_pick(double &best, int &bi, double ob, int oi) {
    if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
}

template <int RPT, int W>
__global__ __launch_bounds__(LU_THREADS) void k_dlu_leaf(DV F, int64_t m, int64_t c0, int w, int64_t *ipiv, Status *st) {
    __shared__ double red_v[LU_WAVES];
    __shared__ int red_i[LU_WAVES];
    __shared__ double prow[W], jrow[W];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int mr = (int)(m - c0);   // <= 32768
    double v[RPT][W];
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
        const int r = t + LU_THREADS * i;
#pragma unroll
        for (int j = 0; j < W; ++j) v[i][j] = (r < mr && j < w) ? F.p[(c0 + r) * F.rs + (c0 + j) * F.cs] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < W; ++j) {
        if (j < w) {                                         // uniform
            double best = -1.0;
            int bi = INT32_MAX;
#pragma unroll
            for (int i = 0; i < RPT; ++i) {
                const int r = t + LU_THREADS * i;
                if (r < mr && r >= j) lu_pick(best, bi, fabs(v[i][j]), r);
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) lu_pick(best, bi, __shfl_xor(best, off), __shfl_xor(bi, off));
            if (lane == 0) { red_v[wave] = best; red_i[wave] = bi; }
            __syncthreads();
            best = red_v[0]; bi = red_i[0];
#pragma unroll
            for (int k = 1; k < LU_WAVES; ++k) lu_pick(best, bi, red_v[k], red_i[k]);
            if (best > 0.0) {                                // uniform
                const int pt = bi % LU_THREADS, pi = bi / LU_THREADS;
                if (t == pt) {
#pragma unroll
                    for (int i = 0; i < RPT; ++i)
                        if (i == pi) {
#pragma unroll
                            for (int jj = 0; jj < W; ++jj) prow[jj] = v[i][jj];
                        }
                }
                if (t == j) {
#pragma unroll
                    for (int jj = 0; jj < W; ++jj) jrow[jj] = v[0][jj];
                }
                __syncthreads();
                if (bi != j) {
                    if (t == j) {
#pragma unroll
                        for (int jj = 0; jj < W; ++jj) v[0][jj] = prow[jj];
                    }
                    if (t == pt) {
#pragma unroll
                        for (int i = 0; i < RPT; ++i)
                            if (i == pi) {
#pragma unroll
                                for (int jj = 0; jj < W; ++jj) v[i][jj] = jrow[jj];
                            }
                    }
                }
                if (t == 0) ipiv[c0 + j] = c0 + bi + 1;
                const double inv = 1.0 / prow[j];
#pragma unroll
                for (int i = 0; i < RPT; ++i) {
                    const int r = t + LU_THREADS * i;
                    if (r < mr && r > j) {
                        const double l = v[i][j] * inv;
                        v[i][j] = l;
#pragma unroll
                        for (int jj = j + 1; jj < W; ++jj) v[i][jj] = fma(-l, prow[jj], v[i][jj]);
                    }
                }
            } else if (t == 0) {                             // zero column
                if (st->info == 0) st->info = c0 + j + 1;
                ipiv[c0 + j] = c0 + j + 1;
            }
            __syncthreads();                                 // LDS reuse
        }
    }
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
        const int r = t + LU_THREADS * i;
        if (r < mr) {
#pragma unroll
            for (int j = 0; j < W; ++j)
                if (j < w) F.p[(c0 + r) * F.rs + (c0 + j) * F.cs] = v[i][j];
        }
    }
}

// k_dlu_swap_solve: one thread per column. Every column of range A
// ([a0, a0 + na)) and range B ([b0, b0 + nb)) gets the row interchanges
// ipiv[k0 .. k0 + kn) in order (dlaswp); range B columns then solve
// L11 x = (P a)(k0 : k0 + kn) with the unit lower triangle L11 =
// F(k0 : k0 + kn, k0 : k0 + kn) by forward substitution (x in registers).
template <int KN>
__global__ __launch_bounds__(256) void k_dlu_swap_solve(DV F, int64_t k0, int kn, const int64_t *ipiv, int64_t a0, int64_t na,
                                                        int64_t b0, int64_t nb) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= na + nb) return;
    const bool solve = e >= na;
    const int64_t j = solve ? b0 + (e - na) : a0 + e;
    double *col = F.p + j * F.cs;
    for (int k = 0; k < kn; ++k) {
        const int64_t r = k0 + k, p = ipiv[r] - 1;
        if (p != r) {
            const double a = col[r * F.rs], b = col[p * F.rs];
            col[r * F.rs] = b;
            col[p * F.rs] = a;
        }
    }
    if (!solve) return;
    double x[KN];
#pragma unroll
    for (int r = 0; r < KN; ++r) {
        if (r < kn) {
            const double *lrow = F.p + (k0 + r) * F.rs + k0 * F.cs;
            double acc = col[(k0 + r) * F.rs];
#pragma unroll
            for (int q = 0; q < r; ++q) acc = fma(-lrow[q * F.cs], x[q], acc);
            x[r] = acc;
            col[(k0 + r) * F.rs] = acc;
        }
    }
}

// ---------------------------------------------------------- launchers
hipError_t dgemm(hipStream_t s, DGemmArgs g, int64_t batch) {
    if (g.m == 0 || g.n == 0 || batch == 0) return hipSuccess;
    if (g.ksplit < 1) g.ksplit = 1;
    if (g.ksplit > 1 && (g.P == nullptr || batch != 1)) return hipErrorInvalidValue;
    const int64_t tiles_m = cdiv(g.m, DG_T), tiles_n = cdiv(g.n, DG_T);
    const dim3 grid((unsigned)(tiles_m * tiles_n), (unsigned)g.ksplit, (unsigned)batch);
    hipLaunchKernelGGL(k_dgemm, grid, dim3(256), 0, s, g, (int)tiles_m);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || g.ksplit == 1) return e;
    hipLaunchKernelGGL(k_dgemm_reduce, dim3((unsigned)cdiv(g.m * g.n, 256)), dim3(256), 0, s, g);
    return hipGetLastError();
}

hipError_t dpotrf_diag(hipStream_t s, const DV &F, const DV &Inv, int64_t j0, int nb, Status *st) {
    if (nb < 1 || nb > 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_dpotrf_diag, dim3(1), dim3(DP_THREADS), 0, s, F, Inv, j0, nb, st);
    return hipGetLastError();
}

hipError_t dcopy(hipStream_t s, int64_t m, int64_t n, const DV &dst, const DV &src, int mirror) {
    if (m * n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_dcopy, dim3((unsigned)cdiv(m * n, 256)), dim3(256), 0, s, m, n, dst, src, mirror);
    return hipGetLastError();
}

hipError_t drow_abs_sum(hipStream_t s, int64_t m, int64_t n, const DV &A, int uplo, double *out, double *part) {
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(k_drow_abs_sum, dim3((unsigned)cdiv(m, 256), ROW_CHUNKS), dim3(256), 0, s, m, n, A, uplo, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_drow_abs_finish, dim3((unsigned)cdiv(m, 256)), dim3(256), 0, s, m, part, out);
    return hipGetLastError();
}

int drow_abs_chunks() { return ROW_CHUNKS; }

int dlu_leaf_width(int64_t m) {
    if (m <= 2048) return 16;
    if (m <= 4096) return 8;
    if (m <= 8192) return 4;
    if (m <= 16384) return 2;
    return m <= 32768 ? 1 : 0;
}

hipError_t dlu_leaf(hipStream_t s, const DV &F, int64_t m, int64_t c0, int w, int64_t *ipiv, Status *st) {
    // the instance is chosen by the rows left (mr <= m, so its width is at
    // least dlu_leaf_width(m))
    const int64_t mr = m - c0;
    if (w < 1 || mr < 1 || w > dlu_leaf_width(mr)) return hipErrorInvalidValue;
    const dim3 g(1), b(LU_THREADS);
    if (mr <= 1024) hipLaunchKernelGGL((k_dlu_leaf<1, 16>), g, b, 0, s, F, m, c0, w, ipiv, st);
    else if (mr <= 2048) hipLaunchKernelGGL((k_dlu_leaf<2, 16>), g, b, 0, s, F, m, c0, w, ipiv, st);
    else if (mr <= 4096) hipLaunchKernelGGL((k_dlu_leaf<4, 8>), g, b, 0, s, F, m, c0, w, ipiv, st);
    else if (mr <= 8192) hipLaunchKernelGGL((k_dlu_leaf<8, 4>), g, b, 0, s, F, m, c0, w, ipiv, st);
    else if (mr <= 16384) hipLaunchKernelGGL((k_dlu_leaf<16, 2>), g, b, 0, s, F, m, c0, w, ipiv, st);
    else hipLaunchKernelGGL((k_dlu_leaf<32, 1>), g, b, 0, s, F, m, c0, w, ipiv, st);
    return hipGetLastError();
}

hipError_t dlu_swap_solve(hipStream_t s, const DV &F, int64_t k0, int kn, const int64_t *ipiv, int64_t a0, int64_t na, int64_t b0,
                          int64_t nb) {
    if (na + nb == 0 || kn == 0) return hipSuccess;
    if (kn < 0 || kn > 64) return hipErrorInvalidValue;
    const dim3 g((unsigned)cdiv(na + nb, 256)), b(256);
    if (kn <= 16) hipLaunchKernelGGL(k_dlu_swap_solve<16>, g, b, 0, s, F, k0, kn, ipiv, a0, na, b0, nb);
    else hipLaunchKernelGGL(k_dlu_swap_solve<64>, g, b, 0, s, F, k0, kn, ipiv, a0, na, b0, nb);
    return hipGetLastError();
}

hipError_t dgather_rows(hipStream_t s, int64_t m, int64_t ncols, const DV &dst, const DV &src, const int64_t *perm) {
    if (m * ncols == 0) return hipSuccess;
    hipLaunchKernelGGL(k_dgather_rows, dim3((unsigned)cdiv(m * ncols, 256)), dim3(256), 0, s, m, ncols, dst, src, perm);
    return hipGetLastError();
}

}  // namespace mfk
