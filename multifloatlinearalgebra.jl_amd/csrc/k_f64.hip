// Binary64 kernels of the refinement's low-precision factor and correction
// solve. These run in plain FP64 (no multi-double networks, no MFLA order
// constraints), so the GEMM-shaped work goes to the CDNA4 FP64 matrix cores:
//
//  * k_dgemm: C = alpha * op(A) * op(B) + beta * C on strided views, 64x64
//    tile per 256-thread workgroup, each wave a 32x32 block of 2x2 MFMA
//    16x16x4 f64 accumulators. Operands pass through double-buffered LDS
//    stages 32 deep; the next stage is fetched into registers while the
//    current one is multiplied.
//    Optional triangular structure of A / B (the k range of each tile is
//    clipped and the zero triangle's values are never used), a
//    lower-triangular output mask, a batch index (grid z) with element
//    strides, and a split-k mode that writes partial sums for k_dgemm_reduce
//    (fixed summation order, so the result is deterministic).
//  * k_dpotrf_diag: Cholesky of one nb x nb (nb <= 64) diagonal block in LDS,
//    plus the inverse of the resulting triangle (needed by the panel update
//    L21 = A21 * inv(L11)^T and by the correction solve).
//  * small helpers: copy of a binary64 plane with optional mirroring, row
//    sums of |A|, row gathers by a permutation.
#include "mf_common.hpp"
#include "k_f64.hpp"

namespace mfk {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int DG_T = 64;

// v_mfma_f64_16x16x4_f64 operand layout (lane l of the wave), as measured
// on gfx950 with integer-valued operands:
//   first operand:  P[l % 16][l / 16]   (16 x 4)
//   second operand: Q[l / 16][l % 16]   (4 x 16)
//   D item v:       D[l / 16 + 4 * v][l % 16]   (D = P Q)
// The kernel forms the transposed block D = B^T A^T (P = B^T, Q = A^T), so
// lane l holds C rows i = l % 16: the 16 lanes of a row group store 16
// consecutive elements of a column-major C column. Inside a 16-deep chunk
// at k0, MFMA step s gives slot q = l / 16 the index k = k0 + 4 q + s, so
// each lane reads four consecutive k of its row of A / column of B.
constexpr int DG_KC = 32, DG_LDK = DG_T + 1;   // k depth of one LDS stage; padded LDS rows

// Operand staging: thread t brings 8 elements of A (64 rows x 32 k) and 8
// of B (32 k x 64 columns) per stage. The lane index runs along whichever
// operand dimension has unit stride, so the loads coalesce for both
// column-major views and transposed ones. Addresses are clamped into the
// matrix and masked elements (outside the range, or in a zero triangle that
// is never read) are replaced by zeros after the load, so no load is
// conditional.
struct DgStage {
    double a[8], b[8];
};

template <typename IT>
__device__ __forceinline__ void dg_fetch(const DGemmArgs &g, const double *Ap, const double *Bp, int64_t i0, int64_t j0, int64_t k0,
                                         int64_t khi, int t, DgStage &st) {
    const bool a_rows = g.A.rs == 1, b_cols = g.B.cs == 1;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int e = t + 256 * q;                                          // 0 .. 2047 = 64 x 32
        const int ii = a_rows ? (e & 63) : (e >> 5), kk = a_rows ? (e >> 6) : (e & 31);
        const int64_t i = i0 + ii, k = k0 + kk;
        bool ok = i < g.m && k < khi;
        if (g.a_tri == TRI_LOWER) ok = ok && k <= i;
        if (g.a_tri == TRI_UPPER) ok = ok && k >= i;
        const IT ic = (IT)(i < g.m ? i : g.m - 1), kc = (IT)(k < g.k ? k : g.k - 1);
        const double v = Ap[ic * (IT)g.A.rs + kc * (IT)g.A.cs];
        st.a[q] = ok ? v : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int e = t + 256 * q;
        const int jj = b_cols ? (e & 63) : (e >> 5), kk = b_cols ? (e >> 6) : (e & 31);
        const int64_t j = j0 + jj, k = k0 + kk;
        bool ok = j < g.n && k < khi;
        if (g.b_tri == TRI_LOWER) ok = ok && k >= j;
        if (g.b_tri == TRI_UPPER) ok = ok && k <= j;
        const IT jc = (IT)(j < g.n ? j : g.n - 1), kc = (IT)(k < g.k ? k : g.k - 1);
        const double v = Bp[kc * (IT)g.B.rs + jc * (IT)g.B.cs];
        st.b[q] = ok ? v : 0.0;
    }
}

__device__ __forceinline__ void dg_store(const DGemmArgs &g, int t, const DgStage &st, double (*As)[DG_LDK], double (*Bs)[DG_LDK]) {
    const bool a_rows = g.A.rs == 1, b_cols = g.B.cs == 1;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int e = t + 256 * q;
        const int ii = a_rows ? (e & 63) : (e >> 5), kk = a_rows ? (e >> 6) : (e & 31);
        As[kk][ii] = st.a[q];
        const int jj = b_cols ? (e & 63) : (e >> 5), k2 = b_cols ? (e >> 6) : (e & 31);
        Bs[k2][jj] = st.b[q];
    }
}

// Inside a stage, MFMA step s of wave (wy, wx) takes k = 4 s + l / 16 for
// lane l from LDS. IT: the integer type of element offsets inside A and B
// (int when the host has checked that every offset fits).
template <typename IT>
__global__ __launch_bounds__(256) void k_dgemm(DGemmArgs g, int tiles_m) {
    __shared__ double As[2][DG_KC][DG_LDK];
    __shared__ double Bs[2][DG_KC][DG_LDK];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wy = w & 1, wx = w >> 1;
    const int64_t tm = blockIdx.x % tiles_m, tn = blockIdx.x / tiles_m;
    const int64_t i0 = tm * DG_T, j0 = tn * DG_T;
    const int64_t z = blockIdx.z;
    const double *Ap = g.A.p + z * g.za, *Bp = g.B.p + z * g.zb;
    double *Cp = g.C.p + z * g.zc;
    if (g.c_lower && j0 > i0 + DG_T - 1) return;  // tile entirely above the diagonal (uniform)

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
        const int64_t chunk = ((span + g.ksplit - 1) / g.ksplit + DG_KC - 1) / DG_KC * DG_KC;
        const int64_t a = klo + ks * chunk, b = a + chunk;
        klo = a;
        khi = b < khi ? b : khi;
    }

    const int64_t ib = i0 + wy * 32 + (lane & 15);
    const int64_t jc = j0 + wx * 32 + (lane >> 4);   // C column of item v of block bj: jc + 16 bj + 4 v

    d4 acc[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = d4{0.0, 0.0, 0.0, 0.0};

    const int64_t nst = khi > klo ? (khi - klo + DG_KC - 1) / DG_KC : 0;
    DgStage st;
    if (nst > 0) {
        dg_fetch<IT>(g, Ap, Bp, i0, j0, klo, khi, t, st);
        dg_store(g, t, st, As[0], Bs[0]);
    }
    __syncthreads();
    for (int64_t s = 0; s < nst; ++s) {
        const int cur = (int)(s & 1);
        if (s + 1 < nst) dg_fetch<IT>(g, Ap, Bp, i0, j0, klo + (s + 1) * DG_KC, khi, t, st);
#pragma unroll
        for (int k4 = 0; k4 < DG_KC / 4; ++k4) {
            const int kr = k4 * 4 + (lane >> 4);
            double a[2], b[2];
#pragma unroll
            for (int bi = 0; bi < 2; ++bi) a[bi] = As[cur][kr][wy * 32 + bi * 16 + (lane & 15)];
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) b[bj] = Bs[cur][kr][wx * 32 + bj * 16 + (lane & 15)];
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[bj], a[bi], acc[bi][bj], 0, 0, 0);
        }
        if (s + 1 < nst) dg_store(g, t, st, As[cur ^ 1], Bs[cur ^ 1]);
        __syncthreads();
    }
    // (C may alias A -- the in-place panel product L21 = A21 inv(L11)^T: all
    // operands of the tile were read before the last barrier above.)
    // Epilogue: the 16 C values are loaded together (clamped addresses, so
    // no load waits on a branch), then the masked stores.
    double cv[2][2][4];
    const bool readc = !g.beta_zero && g.ksplit == 1;
    if (readc) {
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int64_t i = ib + 16 * bi, j = jc + 16 * bj + 4 * v;
                    const int64_t icl = i < g.m ? i : g.m - 1, jcl = j < g.n ? j : g.n - 1;
                    cv[bi][bj][v] = Cp[icl * g.C.rs + jcl * g.C.cs];
                }
    }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t i = ib + 16 * bi, j = jc + 16 * bj + 4 * v;
                const bool ok = i < g.m && j < g.n && !(g.c_lower && i < j);
                const double r = acc[bi][bj][v];
                if (g.ksplit > 1) {
                    if (ok) g.P[((int64_t)ks * g.n + j) * g.m + i] = r;
                } else {
                    const double o = g.beta_zero ? g.alpha * r : g.alpha * r + g.beta * cv[bi][bj][v];
                    if (ok) Cp[i * g.C.rs + j * g.C.cs] = o;
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
// The block is padded to 64 x 64 with an identity and held in LDS by one
// 256-thread workgroup. Factorization in four 16-column panels, three
// barriers each: wave 0 factors the 16 x 16 diagonal block in registers
// (lane i holds row i; the pivot and multipliers reach the other lanes by
// lane reads), one thread per row solves the rows below against it, and the
// workgroup applies the rank-16 update to the trailing lower triangle.
// Inverse: wave 0 inverts the four diagonal 16 x 16 blocks at once (lane
// 16 p + j computes column j of block p), then the off-diagonal blocks by
// block distance d = 1, 2, 3: W = sum_k L(a, k) X(k, b) (kept in the unused
// upper triangle), X(a, b) = -X(a, a) W. A non-positive or non-finite pivot
// at block column c records info = j0 + c + 1 and stops the factorization.
constexpr int DP_LD = 65, DP_THREADS = 256;

__device__ __forceinline__ double dp_bcast(double x, int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), l), __builtin_amdgcn_readlane(__double2loint(x), l));
}

__global__ __launch_bounds__(DP_THREADS) void k_dpotrf_diag(DV F, DV Inv, int64_t j0, int nb, Status *st) {
    if (st->info != 0) return;
    __shared__ double A[64 * DP_LD];   // row-major A[r * DP_LD + q]: the block, L after the factorization
    __shared__ double X[64 * DP_LD];   // row-major inverse (the upper triangle holds temporaries)
    __shared__ double idg[64];         // 1 / L(c, c)
    __shared__ int fail;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    {
        double g[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {                       // clamped addresses: every load is in bounds
            const int e = t + DP_THREADS * k, r = e & 63, q = e >> 6;
            const int rr = r < nb ? r : nb - 1, qq = q < nb ? q : nb - 1;
            g[k] = F.p[(j0 + rr) * F.rs + (j0 + qq) * F.cs];
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = t + DP_THREADS * k, r = e & 63, q = e >> 6;
            const bool in = r < nb && q < nb && r >= q;
            A[r * DP_LD + q] = in ? g[k] : (r == q ? 1.0 : 0.0);
        }
    }
    if (t == 0) fail = 0;
    __syncthreads();

#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const int c0 = 16 * p;
        if (w == 0) {
            // the 16 x 16 diagonal block: lane i (< 16) holds row c0 + i
            const int i = lane & 15;
            double b[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) b[q] = A[(c0 + i) * DP_LD + c0 + q];
            int bad = -1;
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                if (bad < 0) {                                // uniform
                    const double d = dp_bcast(b[k], k);
                    if (!(d > 0.0) || !__builtin_isfinite(d)) {
                        bad = k;
                    } else {
                        const double sd = __builtin_sqrt(d), is = 1.0 / sd;
                        b[k] = i == k ? sd : (i > k ? b[k] * is : b[k]);
#pragma unroll
                        for (int q = k + 1; q < 16; ++q) {
                            const double lq = dp_bcast(b[k], q);
                            b[q] = i >= q ? fma(-b[k], lq, b[q]) : b[q];
                        }
                    }
                }
            }
            if (lane < 16) {
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    if (q <= i) A[(c0 + i) * DP_LD + c0 + q] = b[q];
                    if (q == i) idg[c0 + i] = 1.0 / b[q];
                }
            }
            if (bad >= 0 && lane == 0) {
                fail = 1;
                st->info = j0 + c0 + bad + 1;                 // padding pivots are 1: c0 + bad < nb
            }
        }
        __syncthreads();
        if (fail) return;                                     // uniform
        const int R = 48 - c0;                                // rows below the panel
        if (t < R) {
            const int r = c0 + 16 + t;
            double x[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) x[k] = A[r * DP_LD + c0 + k];
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                double a = x[k];
#pragma unroll
                for (int q = 0; q < k; ++q) a = fma(-x[q], A[(c0 + k) * DP_LD + c0 + q], a);
                x[k] = a * idg[c0 + k];
            }
#pragma unroll
            for (int k = 0; k < 16; ++k) A[r * DP_LD + c0 + k] = x[k];
        }
        __syncthreads();
        for (int e = t; e < R * R; e += DP_THREADS) {
            const int i = c0 + 16 + e % R, j = c0 + 16 + e / R;
            if (i >= j) {
                double acc = A[i * DP_LD + j];
#pragma unroll
                for (int k = 0; k < 16; ++k) acc = fma(-A[i * DP_LD + c0 + k], A[j * DP_LD + c0 + k], acc);
                A[i * DP_LD + j] = acc;
            }
        }
        __syncthreads();
    }

    // ---- inverse: the diagonal 16 x 16 blocks (wave 0, lane 16 p + j)
    if (w == 0) {
        const int c0 = 16 * (lane >> 4), j = lane & 15;
        double x[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            double a = i == j ? 1.0 : 0.0;
#pragma unroll
            for (int k = 0; k < i; ++k) a = fma(-A[(c0 + i) * DP_LD + c0 + k], x[k], a);
            x[i] = i >= j ? a * idg[c0 + i] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) if (i >= j) X[(c0 + i) * DP_LD + c0 + j] = x[i];
    }
    __syncthreads();
    // off-diagonal blocks (a, b), a = b + d
#pragma unroll
    for (int d = 1; d < 4; ++d) {
        const int nblk = 4 - d;
        for (int e = t; e < 256 * nblk; e += DP_THREADS) {   // W(i, jj) = sum_k L(16a + i, k) X(k, 16b + jj), k in [16b, 16a)
            const int bb = e >> 8, i = (e >> 4) & 15, jj = e & 15, aa = bb + d;
            double acc = 0.0;
            for (int k = 16 * bb + jj; k < 16 * aa; ++k) acc = fma(A[(16 * aa + i) * DP_LD + k], X[k * DP_LD + 16 * bb + jj], acc);
            X[(16 * bb + jj) * DP_LD + 16 * aa + i] = acc;   // upper-triangle temporary
        }
        __syncthreads();
        for (int e = t; e < 256 * nblk; e += DP_THREADS) {   // X(16a + i, 16b + jj) = -sum_m X(16a + i, 16a + m) W(m, jj)
            const int bb = e >> 8, i = (e >> 4) & 15, jj = e & 15, aa = bb + d;
            double acc = 0.0;
#pragma unroll
            for (int m = 0; m < 16; ++m)
                if (m <= i) acc = fma(X[(16 * aa + i) * DP_LD + 16 * aa + m], X[(16 * bb + jj) * DP_LD + 16 * aa + m], acc);
            X[(16 * aa + i) * DP_LD + 16 * bb + jj] = -acc;
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int e = t + DP_THREADS * k, r = e & 63, q = e >> 6;
        if (r < nb && q < nb && r >= q) {
            F.p[(j0 + r) * F.rs + (j0 + q) * F.cs] = A[r * DP_LD + q];
            Inv.p[(j0 + r) * Inv.rs + (j0 + q) * Inv.cs] = X[r * DP_LD + q];
        }
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
// rows [c0, m) by one 512-thread workgroup with the panel in registers:
// thread t holds rows c0 + t + 512 i (i < RPT); two waves per SIMD keep the
// per-column instruction stream short. One barrier per column: every wave
// reduces its (|a|, row) candidates with DPP row rotations and lane reads
// (maximum first, then the smallest row holding it), and the lane holding the wave's
// winner publishes that whole row (and the owner of row j publishes row j)
// in a double-buffered LDS slot before the barrier, so after it every
// thread picks the pivot, swaps, scales by the reciprocal pivot and applies
// the rank-1 update of the leaf's later columns without a second exchange.
// ipiv gets 1-based global row indices; a zero (or NaN-only) pivot column
// records info = column + 1 if no failure is recorded yet and is skipped, as
// dgetf2 skips it.
//
// The leaf also brings the other columns of its 64-column block up to date:
// a row id travels with every panel row, so the final permutation of the
// leaf's rows is known; columns [a0, a0 + na) get the row interchanges and
// columns [b0, b0 + nb) get them and the unit lower solve of their top w
// rows (U12 = L11^{-1} P A12). All reads of those columns happen before a
// barrier and all writes after it.
constexpr int LU_THREADS = 512, LU_WAVES = LU_THREADS / 64;

// Workgroup barrier that orders LDS only: global loads issued before it stay
// in flight (a full __syncthreads() waits for them).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int CTRL>
__device__ __forceinline__ double lu_dpp_d(double x) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Every lane of a 16-lane row ends with the row's maximum (quad
// permutations, then rotations by 4 and 8). The values are |a| or -1; a NaN
// never wins (fmax returns the other operand).
__device__ __forceinline__ double lu_row_max(double x) {
    x = fmax(x, lu_dpp_d<0xB1>(x));    // quad_perm [1, 0, 3, 2]
    x = fmax(x, lu_dpp_d<0x4E>(x));    // quad_perm [2, 3, 0, 1]
    x = fmax(x, lu_dpp_d<0x124>(x));   // row_ror 4
    return fmax(x, lu_dpp_d<0x128>(x)); // row_ror 8
}

__device__ __forceinline__ int lu_row_min(int i) {
    i = min(i, __builtin_amdgcn_update_dpp(0, i, 0xB1, 0xf, 0xf, false));
    i = min(i, __builtin_amdgcn_update_dpp(0, i, 0x4E, 0xf, 0xf, false));
    i = min(i, __builtin_amdgcn_update_dpp(0, i, 0x124, 0xf, 0xf, false));
    return min(i, __builtin_amdgcn_update_dpp(0, i, 0x128, 0xf, 0xf, false));
}

__device__ __forceinline__ double lu_readlane(double x, int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), l), __builtin_amdgcn_readlane(__double2loint(x), l));
}

// Wave argmax in two passes: the maximum value, then the smallest index
// among the lanes holding it (LAPACK's first maximum). Every lane ends with
// the result.
__device__ __forceinline__ void lu_wave_argmax(double &best, int &bi) {
    const double r = lu_row_max(best);
    const double m = fmax(fmax(lu_readlane(r, 0), lu_readlane(r, 16)), fmax(lu_readlane(r, 32), lu_readlane(r, 48)));
    const int c = lu_row_min(best == m ? bi : INT32_MAX);
    bi = min(min(__builtin_amdgcn_readlane(c, 0), __builtin_amdgcn_readlane(c, 16)),
             min(__builtin_amdgcn_readlane(c, 32), __builtin_amdgcn_readlane(c, 48)));
    best = m;
}

template <int RPT, int W>
__global__ __launch_bounds__(LU_THREADS) void k_dlu_leaf(DV F, int64_t m, int64_t c0, int w, int64_t *ipiv, Status *st, int64_t a0, int na,
                                                         int64_t b0, int nb) {
    __shared__ double cand[2][LU_WAVES][W], cval[2][LU_WAVES], jrow[2][W];
    __shared__ int cidx[2][LU_WAVES], crid[2][LU_WAVES], jrid[2];
    __shared__ double l11[W][W];
    __shared__ int rtop[W], mpos[W], msrc[W], nmoved;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int mr = (int)(m - c0);   // <= 32768
    double v[RPT][W];
    int rid[RPT];
    bool rowok[RPT];
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
        const int r = t + LU_THREADS * i;
        rid[i] = r;
        rowok[i] = r < mr;
    }
    // column by column, so the first column's loads complete first and its
    // step overlaps the arrival of the later columns
#pragma unroll
    for (int j = 0; j < W; ++j)
#pragma unroll
        for (int i = 0; i < RPT; ++i) v[i][j] = (rowok[i] && j < w) ? F.p[(c0 + t + LU_THREADS * i) * F.rs + (c0 + j) * F.cs] : 0.0;
    if (t == 0) nmoved = 0;
    int piv[W];   // local pivot row of each column (uniform); stored after the loop so that no
                  // global store is outstanding at the column barriers
#pragma unroll
    for (int j = 0; j < W; ++j) {
        if (j < w) {                                         // uniform
            const int par = j & 1;
            double best = -1.0;
            int bi = INT32_MAX;
#pragma unroll
            for (int i = 0; i < RPT; ++i) {            // rows ascending: first maximum kept
                const int r = t + LU_THREADS * i;
                const double a = fabs(v[i][j]);
                // rows of slot i >= 1 are >= 512 > j: only slot 0 compares with j
                const bool take = (i == 0 ? (rowok[0] && t >= j) : rowok[i]) && a > best;
                best = take ? a : best;
                bi = take ? r : bi;
            }
            lu_wave_argmax(best, bi);
#pragma unroll
            for (int i = 0; i < RPT; ++i)
                if (t + LU_THREADS * i == bi) {
#pragma unroll
                    for (int jj = 0; jj < W; ++jj) cand[par][wave][jj] = v[i][jj];
                    crid[par][wave] = rid[i];
                }
            if (lane == 0) { cval[par][wave] = best; cidx[par][wave] = bi; }
            if (t == j) {
#pragma unroll
                for (int jj = 0; jj < W; ++jj) jrow[par][jj] = v[0][jj];
                jrid[par] = rid[0];
            }
            lds_barrier();   // the slots are LDS; the panel's later columns may still be arriving
            // the wave winners: one LDS read per lane of row 0, a row reduction
            {
                const double cb = lane < LU_WAVES ? cval[par][lane] : -1.0;
                const int ci = lane < LU_WAVES ? cidx[par][lane] : INT32_MAX;
                best = lu_readlane(lu_row_max(cb), 0);
                bi = __builtin_amdgcn_readlane(lu_row_min(cb == best ? ci : INT32_MAX), 0);
            }
            const int ww = (bi & (LU_THREADS - 1)) >> 6;     // the wave holding row bi (when best > 0)
            if (best > 0.0) {                                // uniform
                const double *prow = cand[par][ww];
                if (bi != j) {
                    if (t == j) {                            // wave 0 (j < 16)
#pragma unroll
                        for (int jj = 0; jj < W; ++jj) v[0][jj] = prow[jj];
                        rid[0] = crid[par][ww];
                    }
                    if (wave == ww) {                        // uniform: only the pivot row's wave
#pragma unroll
                        for (int i = 0; i < RPT; ++i)
                            if (t + LU_THREADS * i == bi) {
#pragma unroll
                                for (int jj = 0; jj < W; ++jj) v[i][jj] = jrow[par][jj];
                                rid[i] = jrid[par];
                            }
                    }
                }
                piv[j] = bi;
                const double inv = 1.0 / prow[j];
                // multipliers and rank-1 update of the rows below the pivot:
                // slot 0 holds rows t <= j (the pivot row and rows above it),
                // which keep their values; slots >= 1 are all below it (rows
                // past the panel are padding: never compared or stored)
                double l[RPT];
                const bool below0 = t > j;
#pragma unroll
                for (int i = 0; i < RPT; ++i) l[i] = v[i][j] * inv;
                v[0][j] = below0 ? l[0] : v[0][j];
#pragma unroll
                for (int i = 1; i < RPT; ++i) v[i][j] = l[i];
#pragma unroll
                for (int jj = j + 1; jj < W; ++jj) {
                    const double u = prow[jj];
                    v[0][jj] = below0 ? fma(-l[0], u, v[0][jj]) : v[0][jj];
#pragma unroll
                    for (int i = 1; i < RPT; ++i) v[i][jj] = fma(-l[i], u, v[i][jj]);
                }
            } else {                                         // zero column: no interchange
                piv[j] = j;
                if (t == 0 && st->info == 0) st->info = c0 + j + 1;
            }
        }
    }
    __syncthreads();   // the last column's slots are read; nmoved is visible

#pragma unroll
    for (int j = 0; j < W; ++j)
        if (t == j && j < w) ipiv[c0 + j] = c0 + piv[j] + 1;
    // the factored panel
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
        const int r = t + LU_THREADS * i;
        if (r < mr) {
#pragma unroll
            for (int j = 0; j < W; ++j)
                if (j < w) F.p[(c0 + r) * F.rs + (c0 + j) * F.cs] = v[i][j];
        }
    }
    // row permutation of the leaf: top rows and the moved rows below them
    if (t < w) {
        rtop[t] = rid[0];
#pragma unroll
        for (int q = 0; q < W; ++q) l11[t][q] = v[0][q];
    }
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
        const int r = t + LU_THREADS * i;
        if (r < mr && r >= w && rid[i] != r) {           // at most one such row per column
            const int k = atomicAdd(&nmoved, 1);
            mpos[k] = r;
            msrc[k] = rid[i];
        }
    }
    __syncthreads();
    const int nm = nmoved, ncols = na + nb;
    const int items = nm * ncols + w * na;               // <= 2 * 16 * 63 < 4 * LU_THREADS
    double val[4];
    int64_t dst[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int e = t + LU_THREADS * q;
        dst[q] = -1;
        if (e < items) {
            int pos, src, ci;
            if (e < nm * ncols) {
                const int mi = e / ncols;
                ci = e - mi * ncols;
                pos = mpos[mi];
                src = msrc[mi];
            } else {
                const int e2 = e - nm * ncols, k = e2 / na;
                ci = e2 - k * na;
                pos = k;
                src = rtop[k];
            }
            const int64_t col = ci < na ? a0 + ci : b0 + (ci - na);
            val[q] = F.p[(c0 + src) * F.rs + col * F.cs];
            dst[q] = (c0 + pos) * F.rs + col * F.cs;
        }
    }
    double x[W];
    if (t < nb) {
        const int64_t col = b0 + t;
#pragma unroll
        for (int k = 0; k < W; ++k)
            if (k < w) {
                double a = F.p[(c0 + rtop[k]) * F.rs + col * F.cs];
#pragma unroll
                for (int q = 0; q < k; ++q) a = fma(-l11[k][q], x[q], a);
                x[k] = a;
            }
    }
    __syncthreads();   // every read of the other columns is done
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (dst[q] >= 0) F.p[dst[q]] = val[q];
    if (t < nb) {
        const int64_t col = b0 + t;
#pragma unroll
        for (int k = 0; k < W; ++k)
            if (k < w) F.p[(c0 + k) * F.rs + col * F.cs] = x[k];
    }
}

// k_dlu_rows: the row interchanges ipiv[j0 .. j0 + kn) of one block (kn <=
// 64) for the columns left of it ([0, j0)) and right of it ([jend, n)), and
// for the right columns the unit lower solve U12 = L11^{-1} (P A12). One
// 256-thread workgroup per 64-column chunk (left chunks first). Wave 0
// composes the interchange sequence into a table of (position, source row)
// entries: the kn block rows, then every distinct pivot row below the block
// (at most kn). The chunk's source rows are gathered into LDS before any
// write, so the composed permutation is applied in one pass. The solve
// gives each column a quad of lanes, lane g of the quad owning rows g + 4 i
// in registers; x_q reaches the quad by a DPP quad broadcast, so no barrier
// is needed inside the substitution.
constexpr int DR_LD = 65;

template <int G>
__device__ __forceinline__ double quad_bcast(double x) {
    constexpr int ctrl = G | (G << 2) | (G << 4) | (G << 6);
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), ctrl, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), ctrl, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

template <int Q>
__device__ __forceinline__ void dr_load(double (&l)[16], const double *L, int g) {
#pragma unroll
    for (int i = Q >> 2; i < 16; ++i) l[i] = L[Q * 64 + g + 4 * i];
}

template <int Q>
__device__ __forceinline__ void dr_step(double (&x)[16], const double (&l)[16], int g) {
    // x_Q lives in lane Q % 4 of the quad, slot Q / 4
    const double xq = quad_bcast<Q & 3>(x[Q >> 2]);
    {
        constexpr int i = Q >> 2;
        if (g > (Q & 3)) x[i] = fma(-l[i], xq, x[i]);
    }
#pragma unroll
    for (int i = (Q >> 2) + 1; i < 16; ++i) x[i] = fma(-l[i], xq, x[i]);
}

// steps Q .. kn - 1; lc holds column Q of L, column Q + 1 is loaded while
// step Q runs
template <int Q>
__device__ __forceinline__ void dr_solve(double (&x)[16], double (&lc)[16], const double *L, int g, int kn) {
    if constexpr (Q < 64) {
        if (Q < kn) {
            double ln[16];
            if constexpr (Q + 1 < 64) dr_load<Q + 1>(ln, L, g);
            dr_step<Q>(x, lc, g);
            if constexpr (Q + 1 < 64) dr_solve<Q + 1>(x, ln, L, g, kn);
        }
    }
}

__global__ __launch_bounds__(256) void k_dlu_rows(DV F, int64_t n, int64_t j0, int kn, const int64_t *ipiv, int left_chunks) {
    __shared__ double T[64 * DR_LD];   // T[k * DR_LD + c]: block row j0 + k as it is now
    __shared__ double X[64 * DR_LD];   // X[r * DR_LD + c]: moved row xpos[r] as it is now
    __shared__ double L[64 * 64];      // L[q * 64 + k] = L11(k, q), zero outside the block
    __shared__ int sp[64], tslot[64], xslot[64], xpos[64], nx;
    const int t = threadIdx.x, lane = t & 63;
    const int64_t jend = j0 + kn;
    const bool right = (int)blockIdx.x >= left_chunks;
    const int64_t cbeg = right ? jend + 64 * (int64_t)(blockIdx.x - left_chunks) : 64 * (int64_t)blockIdx.x;
    const int64_t clim = right ? n : j0;
    const int nc = (int)((clim - cbeg) < 64 ? (clim - cbeg) : 64);

    // ---- the pivot rows below the block, each once (wave 0)
    int p = -1, slot = 0;
    if (t < 64) {
        p = lane < kn ? (int)(ipiv[j0 + lane] - 1) : -1;     // global 0-based pivot row of step lane
        sp[lane] = p;
    }
    __syncthreads();
    if (t < 64) {
        const bool extra = lane < kn && p >= (int)jend;
        int first = lane;                                     // first step with the same pivot row
#pragma unroll
        for (int q = 0; q < 64; ++q) {
            const int pq = sp[q];
            first = (q < first && pq == p) ? q : first;
        }
        const bool isfirst = extra && first == lane;
        const unsigned long long fm = __ballot(isfirst);
        const int rank = __popcll(fm & ((1ull << lane) - 1ull));
        const int rfirst = __shfl(rank, first);
        slot = lane < kn ? (extra ? 64 + rfirst : p - (int)j0) : lane;   // table slot of step lane's pivot row
        if (isfirst) xpos[rank] = p;
        if (lane == 0) nx = __popcll(fm);
    }
    __syncthreads();
    const int nxr = nx;
    // ---- gather the chunk's rows by position (block rows, moved rows) and
    // L11; the loads are in flight while wave 0 composes the interchanges
    const int gk = t & 63, gc = t >> 6;
    double g1[16], g2[16], g3[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int c = gc + 4 * q;
        g1[q] = (gk < kn && c < nc) ? F.p[(j0 + gk) * F.rs + (cbeg + c) * F.cs] : 0.0;
        g2[q] = (gk < nxr && c < nc) ? F.p[(int64_t)xpos[gk] * F.rs + (cbeg + c) * F.cs] : 0.0;
        g3[q] = (right && gk < kn && c < kn) ? F.p[(j0 + gk) * F.rs + (j0 + c) * F.cs] : 0.0;
    }
    if (t < 64) {
        // table entries: slot e < 64 is block row j0 + e, slot 64 + r is
        // moved row xpos[r]; sa / sb hold the slot whose current row ends
        // at each entry after the interchange sequence
        int sa = lane, sb = 64 + lane;
#pragma unroll
        for (int k = 0; k < 64; ++k) {
            if (k < kn) {
                const int s = __builtin_amdgcn_readlane(slot, k);
                const int sl = s & 63;
                const bool hb = s >= 64;
                const int a = __builtin_amdgcn_readlane(sa, k);
                const int b = hb ? __builtin_amdgcn_readlane(sb, sl) : __builtin_amdgcn_readlane(sa, sl);
                sa = lane == k ? b : sa;
                sa = (!hb && lane == sl) ? a : sa;
                sb = (hb && lane == sl) ? a : sb;
            }
        }
        tslot[lane] = sa;
        xslot[lane] = sb;
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        T[gk * DR_LD + gc + 4 * q] = g1[q];
        X[gk * DR_LD + gc + 4 * q] = g2[q];
        L[(gc + 4 * q) * 64 + gk] = g3[q];
    }
    __syncthreads();
    // the value at entry e after the interchanges: the row in slot tslot / xslot
    auto row_of = [&](int sl, int c) -> double { return sl < 64 ? T[sl * DR_LD + c] : X[(sl - 64) * DR_LD + c]; };
    if (right) {
        const int c = t >> 2, g = t & 3;
        double x[16], l0[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = g + 4 * i < kn ? row_of(tslot[g + 4 * i], c) : 0.0;
        dr_load<0>(l0, L, g);
        dr_solve<0>(x, l0, L, g, kn);
        // the moved rows are read from X before any of them is overwritten
        double mv[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int cc = gc + 4 * q;
            mv[q] = gk < nxr ? row_of(xslot[gk], cc) : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; ++i) T[(g + 4 * i) * DR_LD + c] = x[i];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int cc = gc + 4 * q;
            if (gk < kn && cc < nc) F.p[(j0 + gk) * F.rs + (cbeg + cc) * F.cs] = T[gk * DR_LD + cc];
            if (gk < nxr && cc < nc && xslot[gk] != 64 + gk) F.p[(int64_t)xpos[gk] * F.rs + (cbeg + cc) * F.cs] = mv[q];
        }
    } else {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int cc = gc + 4 * q;
            if (gk < kn && cc < nc && tslot[gk] != gk) F.p[(j0 + gk) * F.rs + (cbeg + cc) * F.cs] = row_of(tslot[gk], cc);
            if (gk < nxr && cc < nc && xslot[gk] != 64 + gk) F.p[(int64_t)xpos[gk] * F.rs + (cbeg + cc) * F.cs] = row_of(xslot[gk], cc);
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
    // element offsets inside A and B fit in 32 bits?
    auto span = [](int64_t r, int64_t c, const DV &v) { return (r > 0 ? r - 1 : 0) * (v.rs < 0 ? -v.rs : v.rs) + (c > 0 ? c - 1 : 0) * (v.cs < 0 ? -v.cs : v.cs); };
    const bool off32 = span(g.m, g.k, g.A) < INT32_MAX && span(g.k, g.n, g.B) < INT32_MAX;
    if (off32) hipLaunchKernelGGL(k_dgemm<int>, grid, dim3(256), 0, s, g, (int)tiles_m);
    else hipLaunchKernelGGL(k_dgemm<int64_t>, grid, dim3(256), 0, s, g, (int)tiles_m);
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

hipError_t dlu_leaf(hipStream_t s, const DV &F, int64_t m, int64_t c0, int w, int64_t *ipiv, Status *st, int64_t a0, int na, int64_t b0,
                    int nb) {
    // the instance is chosen by the rows left (mr <= m, so its width is at
    // least dlu_leaf_width(m))
    const int64_t mr = m - c0;
    if (w < 1 || mr < w || w > dlu_leaf_width(mr) || na < 0 || nb < 0 || na + nb + w > 64) return hipErrorInvalidValue;
    const dim3 g(1), b(LU_THREADS);
    if (mr <= 1024) hipLaunchKernelGGL((k_dlu_leaf<2, 16>), g, b, 0, s, F, m, c0, w, ipiv, st, a0, na, b0, nb);
    else if (mr <= 2048) hipLaunchKernelGGL((k_dlu_leaf<4, 16>), g, b, 0, s, F, m, c0, w, ipiv, st, a0, na, b0, nb);
    else if (mr <= 4096) hipLaunchKernelGGL((k_dlu_leaf<8, 8>), g, b, 0, s, F, m, c0, w, ipiv, st, a0, na, b0, nb);
    else if (mr <= 8192) hipLaunchKernelGGL((k_dlu_leaf<16, 4>), g, b, 0, s, F, m, c0, w, ipiv, st, a0, na, b0, nb);
    else if (mr <= 16384) hipLaunchKernelGGL((k_dlu_leaf<32, 2>), g, b, 0, s, F, m, c0, w, ipiv, st, a0, na, b0, nb);
    else hipLaunchKernelGGL((k_dlu_leaf<64, 1>), g, b, 0, s, F, m, c0, w, ipiv, st, a0, na, b0, nb);
    return hipGetLastError();
}

hipError_t dlu_rows(hipStream_t s, const DV &F, int64_t n, int64_t j0, int kn, const int64_t *ipiv) {
    if (kn < 1 || kn > 64 || j0 + kn > n) return hipErrorInvalidValue;
    const int64_t left = cdiv(j0, 64), rightc = cdiv(n - j0 - kn, 64);
    if (left + rightc == 0) return hipSuccess;
    hipLaunchKernelGGL(k_dlu_rows, dim3((unsigned)(left + rightc)), dim3(256), 0, s, F, n, j0, kn, ipiv, (int)left);
    return hipGetLastError();
}

hipError_t dgather_rows(hipStream_t s, int64_t m, int64_t ncols, const DV &dst, const DV &src, const int64_t *perm) {
    if (m * ncols == 0) return hipSuccess;
    hipLaunchKernelGGL(k_dgather_rows, dim3((unsigned)cdiv(m * ncols, 256)), dim3(256), 0, s, m, ncols, dst, src, perm);
    return hipGetLastError();
}

}  // namespace mfk
