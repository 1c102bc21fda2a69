// Ozaki-split products on the INT8 matrix cores (v_mfma_i32_16x16x64_i8).
//
// A Float64xN operand row (or column) is scaled by a power of two 2^e with
// |value| < 2^(e-2) over the row, converted exactly (up to truncation below
// 2^(e - 7S)) to a two's complement fixed-point integer Z of 7S bits, and
// split into S balanced base-128 digits d_s in [-64, 63] (most significant
// first):   value ~ 2^(e - 7S) * Z,   Z = sum_s d_s 128^(S-1-s).
// Then
//   sum_k A(i,k) B(k,j) = sum_{s,t} 2^(ea_i + eb_j - 7(s+t+2)) C_{s,t}(i,j),
//   C_{s,t}(i,j) = sum_k dA_s(i,k) dB_t(j,k)          (exact int32)
// Digit products are at most 2^12 in magnitude, so a group of P pairs over K
// terms stays below 2^31 while P*K <= 2^18.5 (K <= 8192 with P <= 34).
//
//  * k_oz_exp_part / k_oz_exp_finish: e_i = ilogb(max_k |hi(X(i,k))|) + 3
//    (chunked max, fixed order), or a sentinel for all-zero rows.
//  * k_oz_split<N>: the exact fixed-point conversion (each limb's 53-bit
//    mantissa shifted into a 256-bit two's complement accumulator, truncated
//    toward zero) and the balanced digit extraction.
//  * k_oz_round_fp64: a binary64 correction rounded to 7T-bit fixed point per
//    column (the rounded value is written back, so the caller uses exactly
//    the value whose digits are multiplied) and split into T digits.
//  * k_oz_gemm: for every u in [0, U], C_u(i,j) = sum_{s+t=u} sum_k dA_s dB_t
//    as int32 planes; 128 x 128 (or 128 x 64 for narrow outputs) tile per
//    256-thread workgroup, each wave a 4 x 4 (4 x 2) block of 16 x 16 MFMA
//    tiles, 128-byte k-stages prefetched into registers and staged through
//    padded LDS rows; grid z = u.
//  * k_oz_combine<N>: acc(i,j) = sum_u C_u(i,j) 2^(ea_i + eb_j - 7(u+2)) in
//    Float64xN arithmetic, u ascending, then C = alpha*acc + beta*C (or
//    R = R - acc for a residual update).
#include <climits>
#include "mf_common.hpp"
#include "k_ozaki.hpp"

namespace mfk {

constexpr int OZ_CHUNKS = 32;
constexpr int OZ_ZERO_ROW = INT_MIN / 2;

// ------------------------------------------------------------ exponents
// view X(i, k): rows i in [0, m), k in [0, K); part[c * m + i] = max over
// chunk c of |hi(X(i, k))|.
// uplo 1 / 2: X is symmetric and only its lower / upper triangle is read.
__device__ __forceinline__ const double *oz_at(const DV &X, int uplo, int64_t i, int64_t k) {
    if ((uplo == 1 && i < k) || (uplo == 2 && i > k)) { const int64_t q = i; i = k; k = q; }
    return X.p + i * X.rs + k * X.cs;
}

__global__ void k_oz_exp_part(int64_t m, int64_t K, DV X, int uplo, double *part) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int64_t c = blockIdx.y, per = (K + OZ_CHUNKS - 1) / OZ_CHUNKS;
    const int64_t klo = c * per, khi = (klo + per < K) ? klo + per : K;
    double mx = 0.0;
    for (int64_t k = klo; k < khi; ++k) mx = fmax(mx, fabs(*oz_at(X, uplo, i, k)));
    part[c * m + i] = mx;
}

__global__ void k_oz_exp_finish(int64_t m, const double *part, int *e) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double mx = 0.0;
    for (int c = 0; c < OZ_CHUNKS; ++c) mx = fmax(mx, part[c * m + i]);
    e[i] = (mx > 0.0 && __builtin_isfinite(mx)) ? ilogb(mx) + 3 : OZ_ZERO_ROW;
}

// ------------------------------------------------------ fixed point
// 256-bit two's complement accumulator, little-endian 64-bit words.
struct Acc256 {
    uint64_t w[4];
};

__device__ __forceinline__ void acc_add_shifted(Acc256 &a, uint64_t mag, int sh, bool neg) {
    // value = (neg ? -1 : 1) * floor-toward-zero(mag * 2^sh), |result| < 2^255
    uint64_t t[4] = {0, 0, 0, 0};
    if (sh >= 0) {
        const int wq = sh >> 6, bq = sh & 63;
        if (wq < 4) {
            t[wq] = mag << bq;
            if (bq != 0 && wq + 1 < 4) t[wq + 1] = mag >> (64 - bq);
        }
    } else {
        const int r = -sh;
        t[0] = r >= 64 ? 0 : (mag >> r);
    }
    if (neg) {   // two's complement negation of t
        uint64_t carry = 1;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint64_t v = ~t[q] + carry;
            carry = (carry && v == 0) ? 1 : 0;
            t[q] = v;
        }
    }
    uint64_t carry = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint64_t s1 = a.w[q] + t[q];
        const uint64_t c1 = s1 < a.w[q] ? 1 : 0;
        const uint64_t s2 = s1 + carry;
        const uint64_t c2 = s2 < s1 ? 1 : 0;
        a.w[q] = s2;
        carry = c1 | c2;
    }
}

// digits of a 7S-bit value held in a 256-bit accumulator: balanced base 128,
// least significant first into d[S-1], ..., most significant into d[0]
__device__ __forceinline__ void acc_digits(const Acc256 &a, int S, int8_t *d, int64_t stride) {
    int carry = 0;
    for (int q = 0; q < S; ++q) {
        const int bit = 7 * q;
        const int wq = bit >> 6, bq = bit & 63;
        uint64_t chunk = a.w[wq] >> bq;
        if (bq > 57 && wq + 1 < 4) chunk |= a.w[wq + 1] << (64 - bq);
        int b = (int)(chunk & 127) + carry;
        int dig;
        if (b >= 64) { dig = b - 128; carry = 1; } else { dig = b; carry = 0; }
        d[(int64_t)(S - 1 - q) * stride] = (int8_t)dig;
    }
}

// acc_digits with the digit count a compile-time constant: every word index
// and shift is static (the runtime form selects among the four words per
// digit).
template <int S>
__device__ __forceinline__ void acc_digits_t(const Acc256 &a, int8_t *d, int64_t stride) {
    int carry = 0;
#pragma unroll
    for (int q = 0; q < S; ++q) {
        const int bit = 7 * q;
        const int wq = bit >> 6, bq = bit & 63;
        uint64_t chunk = a.w[wq] >> bq;
        if (bq > 57 && wq + 1 < 4) chunk |= a.w[wq + 1] << (64 - bq);
        int b = (int)(chunk & 127) + carry;
        int dig;
        if (b >= 64) { dig = b - 128; carry = 1; } else { dig = b; carry = 0; }
        d[(int64_t)(S - 1 - q) * stride] = (int8_t)dig;
    }
}

// Byte offset of digit (i, k) inside one digit plane. Row layout: i * ld +
// k. Fragment layout (FRAG): 16-row blocks, each a sequence of 64-deep k
// steps of 1 KB laid out exactly as the lanes of a wave read one MFMA
// 16x16x64 fragment (lane l = 16 * (k % 64 / 16) + i % 16 takes 16
// consecutive k), so a wave's fragment load is one contiguous 1 KB block.
template <bool FRAG>
__device__ __forceinline__ int64_t oz_off(int64_t i, int64_t k, int64_t ld) {
    if (!FRAG) return i * ld + k;
    return ((i >> 4) * (ld >> 6) + (k >> 6)) * 1024 + ((((k & 63) >> 4) << 4) + (i & 15)) * 16 + (k & 15);
}

// dig[s * plane + oz_off(i, k)] for X(i, k), s = 0..S-1; the padding
// columns k in [K, ld) get zero digits.
template <int N, bool FRAG, int SC>
__global__ void k_oz_split(int64_t m, int64_t K, DV X, int uplo, const int *e, int S, int8_t *dig, int64_t ld, int64_t plane) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= m * ld) return;
    const int64_t k = idx % ld, i = idx / ld;        // k fastest: coalesced digit stores
    int8_t *d = dig + oz_off<FRAG>(i, k, ld);
    const int ei = e[i];
    Acc256 a{{0, 0, 0, 0}};
    if (ei != OZ_ZERO_ROW && k < K) {
        const double *p = oz_at(X, uplo, i, k);
#pragma unroll
        for (int l = 0; l < N; ++l) {
            const double x = p[l * X.ls];
            if (x == 0.0 || !__builtin_isfinite(x)) continue;
            int ex;
            const double f = frexp(fabs(x), &ex);                 // |x| = f 2^ex, f in [0.5, 1)
            const uint64_t mag = (uint64_t)ldexp(f, 53);            // 53-bit integer mantissa
            acc_add_shifted(a, mag, ex - 53 + 7 * (SC > 0 ? SC : S) - ei, x < 0.0);
        }
    }
    if constexpr (SC > 0) acc_digits_t<SC>(a, d, plane);
    else acc_digits(a, S, d, plane);
}

// binary64 correction D(k, j) (k in [0, K), column j): per-column exponent
// e_j from the column max; D(k, j) <- round(D 2^(7T - e_j)) 2^(e_j - 7T), and
// its T digits into dig[t * plane + j * ld + k].
__global__ void k_oz_round_fp64(int64_t K, int64_t ncols, DV D, const int *e, int T, int8_t *dig, int64_t ld, int64_t plane) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ld * ncols) return;
    const int64_t k = idx % ld, j = idx / ld;
    if (k >= K) {                                   // zero padding digits
        for (int t = 0; t < T; ++t) dig[(int64_t)t * plane + j * ld + k] = 0;
        return;
    }
    double *p = D.p + k * D.rs + j * D.cs;
    const int ej = e[j];
    int64_t q = 0;
    if (ej != OZ_ZERO_ROW && __builtin_isfinite(*p)) {
        q = (int64_t)rint(ldexp(*p, 7 * T - ej));                 // |q| < 2^(7T - 2)
        *p = ldexp((double)q, ej - 7 * T);
    } else {
        *p = 0.0;
    }
    int carry = 0;
    const uint64_t u = (uint64_t)q;
    for (int t = 0; t < T; ++t) {
        int b = (int)((u >> (7 * t)) & 127) + carry;
        int dg;
        if (b >= 64) { dg = b - 128; carry = 1; } else { dg = b; carry = 0; }
        dig[(int64_t)(T - 1 - t) * plane + j * ld + k] = (int8_t)dg;
    }
}

// ----------------------------------------------------------- INT8 GEMM
typedef int v4i __attribute__((ext_vector_type(4)));

// out[u * m * n + j * m + i] = sum_{s + t = u, s < SA, t < SB} sum_k dA_s(i,k) dB_t(j,k)
// K is the padded length of the digit rows (a multiple of OZ_KS); padding
// digits are zero. Workgroup tile (32 WM) x (32 WN) from 2 x 2 waves of
// WM x WN MFMA 16x16x64 blocks; every (plane s, 128-byte k-stage) is one
// stage of the loop, with the next stage prefetched into registers while
// the current one is multiplied from LDS. LDS rows are padded to 144 bytes
// so the 16 rows a lane group reads fall on distinct banks.
constexpr int OZ_KS = 128, OZ_LDW = 144;
constexpr int OZ_NARROW_TB = 6;   // B digits of the narrow kernel (binary64 corrections)

template <int WM, int WN>
__global__ __launch_bounds__(256) void k_oz_gemm(int64_t m, int64_t n, int64_t K, const int8_t *dA, int64_t lda, int64_t planeA,
                                                 int SA, const int8_t *dB, int64_t ldb, int64_t planeB, int SB, int32_t *out,
                                                 int tiles_m) {
    constexpr int TM = 32 * WM, TN = 32 * WN;
    constexpr int CA = TM * (OZ_KS / 16) / 256, CB = TN * (OZ_KS / 16) / 256;   // 16-byte chunks per thread
    __shared__ __align__(16) int8_t As[TM * OZ_LDW];
    __shared__ __align__(16) int8_t Bs[TN * OZ_LDW];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wy = w & 1, wx = w >> 1;
    const int64_t i0 = (int64_t)(blockIdx.x % tiles_m) * TM, j0 = (int64_t)(blockIdx.x / tiles_m) * TN;
    const int u = (int)blockIdx.z;
    const int slo = u - (SB - 1) > 0 ? u - (SB - 1) : 0, shi = u < SA - 1 ? u : SA - 1;
    const int64_t kst = K / OZ_KS;
    const int64_t stages = (int64_t)(shi - slo + 1) * kst;
    v4i acc[WM][WN];
#pragma unroll
    for (int a = 0; a < WM; ++a)
#pragma unroll
        for (int b = 0; b < WN; ++b) acc[a][b] = v4i{0, 0, 0, 0};
    v4i ra[CA], rb[CB];
    auto fetch = [&](int64_t g) {
        const int sp = slo + (int)(g / kst);
        const int64_t k0 = (g % kst) * OZ_KS;
        const int8_t *pa = dA + (int64_t)sp * planeA + k0, *pb = dB + (int64_t)(u - sp) * planeB + k0;
#pragma unroll
        for (int q = 0; q < CA; ++q) {
            const int c = t + 256 * q, r = c / (OZ_KS / 16), col = (c % (OZ_KS / 16)) * 16;
            ra[q] = (i0 + r < m) ? *(const v4i *)(pa + (i0 + r) * lda + col) : v4i{0, 0, 0, 0};
        }
#pragma unroll
        for (int q = 0; q < CB; ++q) {
            const int c = t + 256 * q, r = c / (OZ_KS / 16), col = (c % (OZ_KS / 16)) * 16;
            rb[q] = (j0 + r < n) ? *(const v4i *)(pb + (j0 + r) * ldb + col) : v4i{0, 0, 0, 0};
        }
    };
    if (stages > 0) fetch(0);
    for (int64_t g = 0; g < stages; ++g) {
#pragma unroll
        for (int q = 0; q < CA; ++q) {
            const int c = t + 256 * q, r = c / (OZ_KS / 16), col = (c % (OZ_KS / 16)) * 16;
            *(v4i *)(As + r * OZ_LDW + col) = ra[q];
        }
#pragma unroll
        for (int q = 0; q < CB; ++q) {
            const int c = t + 256 * q, r = c / (OZ_KS / 16), col = (c % (OZ_KS / 16)) * 16;
            *(v4i *)(Bs + r * OZ_LDW + col) = rb[q];
        }
        __syncthreads();
        if (g + 1 < stages) fetch(g + 1);
#pragma unroll
        for (int ks = 0; ks < OZ_KS / 64; ++ks) {
            v4i fa[WM], fb[WN];
#pragma unroll
            for (int a = 0; a < WM; ++a) fa[a] = *(const v4i *)(As + (wy * 16 * WM + a * 16 + (lane & 15)) * OZ_LDW + ks * 64 + (lane >> 4) * 16);
#pragma unroll
            for (int b = 0; b < WN; ++b) fb[b] = *(const v4i *)(Bs + (wx * 16 * WN + b * 16 + (lane & 15)) * OZ_LDW + ks * 64 + (lane >> 4) * 16);
#pragma unroll
            for (int a = 0; a < WM; ++a)
#pragma unroll
                for (int b = 0; b < WN; ++b) acc[a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fb[b], fa[a], acc[a][b], 0, 0, 0);
        }
        __syncthreads();
    }
    // with the B fragment first, item v of lane l holds row (lane % 16) and
    // column 4 (lane / 16) + v of the 16 x 16 block: coalesced column stores
    int32_t *o = out + (int64_t)u * m * n;
#pragma unroll
    for (int a = 0; a < WM; ++a)
#pragma unroll
        for (int b = 0; b < WN; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t i = i0 + wy * 16 * WM + a * 16 + (lane & 15);
                const int64_t j = j0 + wx * 16 * WN + b * 16 + 4 * (lane >> 4) + v;
                if (i < m && j < n) o[j * m + i] = acc[a][b][v];
            }
}

// Narrow products (the refinement's A * D: few columns, TB = 6 digits of a
// binary64 correction): every wave owns one 16 x 16 output block and keeps
// the partial sums of all SMAX + TB - 1 planes in registers, so each digit
// plane of A is streamed once per 64-column group. A's digits are in the
// fragment layout (one contiguous 1 KB load per wave and 64-deep k step);
// the TB digit rows of the wave's 16 columns are few and stay in L2. Per k
// step the wave loads its TB B fragments once and multiplies every A digit
// s into the planes s .. s + TB - 1. A workgroup is 16 rows x 64 columns
// (four waves side by side); grid.y runs over 64-column groups.
template <int SA, int TB>
__global__ __launch_bounds__(256) void k_oz_gemm_narrow(int64_t m, int64_t n, int64_t K, const int8_t *dA, int64_t planeA,
                                                        const int8_t *dB, int64_t ldb, int64_t planeB, int32_t *out) {
    constexpr int NP = SA + TB - 1;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t i0 = (int64_t)blockIdx.x * 16, j0 = (int64_t)blockIdx.y * 64 + (int64_t)w * 16;
    const int64_t ia = i0 + (lane & 15), jb = j0 + (lane & 15);
    const bool rowok = ia < m, colok = jb < n;
    // the fragment layout pads the rows to 16, so every A load is in bounds
    const int8_t *pa = dA + (int64_t)blockIdx.x * (K >> 6) * 1024 + lane * 16;
    const int8_t *pb = dB + (colok ? jb : 0) * ldb + (lane >> 4) * 16;
    v4i acc[NP];
#pragma unroll
    for (int u = 0; u < NP; ++u) acc[u] = v4i{0, 0, 0, 0};
    // all A fragments of a k step in registers; each slot is reloaded for
    // the next step right after its MFMAs are issued
    v4i fa[SA], fb[TB];
#pragma unroll
    for (int s = 0; s < SA; ++s) fa[s] = *(const v4i *)(pa + s * planeA);
#pragma unroll
    for (int q = 0; q < TB; ++q) fb[q] = *(const v4i *)(pb + q * planeB);
    const v4i zero{0, 0, 0, 0};
    for (int64_t k0 = 0; k0 < K; k0 += 64) {
        const bool more = k0 + 64 < K;
        const int8_t *pn = pa + ((k0 + 64) >> 6) * 1024;
        v4i gb[TB];
#pragma unroll
        for (int q = 0; q < TB; ++q) gb[q] = colok ? fb[q] : zero;
        if (more) {
#pragma unroll
            for (int q = 0; q < TB; ++q) fb[q] = *(const v4i *)(pb + q * planeB + k0 + 64);
        }
#pragma unroll
        for (int s = 0; s < SA; ++s) {
            const v4i a = fa[s];
#pragma unroll
            for (int q = 0; q < TB; ++q) acc[s + q] = __builtin_amdgcn_mfma_i32_16x16x64_i8(gb[q], a, acc[s + q], 0, 0, 0);
            if (more) fa[s] = *(const v4i *)(pn + s * planeA);
        }
    }
    // item v of lane l: row (lane % 16), column 4 (lane / 16) + v of the block
#pragma unroll
    for (int u = 0; u < NP; ++u) {
        int32_t *o = out + (int64_t)u * m * n;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t j = j0 + 4 * (lane >> 4) + v;
            if (rowok && j < n) o[j * m + ia] = acc[u][v];
        }
    }
}

// -------------------------------------------------------------- combine
// acc = sum_u C_u 2^(ea_i + eb_j - 7(u + 2)), u ascending, in N limbs; then
// mode 0: C = alpha*acc (beta == 0) or alpha*acc + beta*C; mode 1: C = C - acc.
template <int N>
__global__ void k_oz_combine(int64_t m, int64_t n, int U, const int32_t *planes, const int *ea, const int *eb, DV C, S<N> alpha,
                             S<N> beta, int beta_zero, int mode) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= m * n) return;
    const int64_t i = idx % m, j = idx / m;
    mf::V<N> acc = mf::zero<N>();
    if (ea[i] != OZ_ZERO_ROW && eb[j] != OZ_ZERO_ROW) {
        const int base = ea[i] + eb[j];
        for (int u = 0; u <= U; ++u) {
            const int32_t c = planes[(int64_t)u * m * n + j * m + i];
            if (c != 0) acc = mf::add(acc, mf::from<N>(ldexp((double)c, base - 7 * (u + 2))));
        }
    }
    mf::V<N> r;
    if (mode == 1) r = mf::sub(ld<N>(C, i, j), acc);
    else if (beta_zero) r = mf::mul(sv(alpha), acc);
    else r = mf::add(mf::mul(sv(alpha), acc), mf::mul(sv(beta), ld<N>(C, i, j)));
    stv<N>(C, i, j, r);
}

// ---------------------------------------------------------- launchers
#define OZ_DISPATCH(N, ...)                                        \
    switch (N) {                                                   \
        case 1: { constexpr int NN = 1; __VA_ARGS__; } break;      \
        case 2: { constexpr int NN = 2; __VA_ARGS__; } break;      \
        case 3: { constexpr int NN = 3; __VA_ARGS__; } break;      \
        case 4: { constexpr int NN = 4; __VA_ARGS__; } break;      \
        default: return hipErrorInvalidValue;                      \
    }

template <int N> static S<N> mk_oz(const double *a) { S<N> s; for (int l = 0; l < N; ++l) s.x[l] = a ? a[l] : 0.0; return s; }

int oz_digits(int N) { return N >= 4 ? 34 : (N == 3 ? 27 : (N == 2 ? 19 : 11)); }
int oz_chunks() { return OZ_CHUNKS; }

hipError_t oz_exponents(hipStream_t s, int64_t m, int64_t K, const DV &X, int uplo, int *e, double *part) {
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(k_oz_exp_part, dim3((unsigned)cdiv(m, 256), OZ_CHUNKS), dim3(256), 0, s, m, K, X, uplo, part);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_oz_exp_finish, dim3((unsigned)cdiv(m, 256)), dim3(256), 0, s, m, part, e);
    return hipGetLastError();
}

hipError_t oz_split(int N, hipStream_t s, int64_t m, int64_t K, const DV &X, int uplo, const int *e, int S, int8_t *dig, int64_t ld,
                    int64_t plane, bool frag) {
    if (m * ld == 0) return hipSuccess;
    if (S < 1 || 7 * S > 250 || ld % 64 != 0 || ld < K) return hipErrorInvalidValue;
    if (frag && plane < cdiv(m, 16) * 16 * ld) return hipErrorInvalidValue;
    const dim3 grid((unsigned)cdiv(m * ld, 256));
#define OZ_SPLIT(F, SCV) OZ_DISPATCH(N, hipLaunchKernelGGL((k_oz_split<NN, F, SCV>), grid, dim3(256), 0, s, m, K, X, uplo, e, S, dig, ld, plane))
#define OZ_SPLIT_S(F)                                            \
    switch (S) {                                                 \
        case 11: OZ_SPLIT(F, 11); break;                         \
        case 19: OZ_SPLIT(F, 19); break;                         \
        case 27: OZ_SPLIT(F, 27); break;                         \
        case 34: OZ_SPLIT(F, 34); break;                         \
        default: OZ_SPLIT(F, 0); break;                          \
    }
    if (frag) { OZ_SPLIT_S(true) } else { OZ_SPLIT_S(false) }
#undef OZ_SPLIT_S
#undef OZ_SPLIT
    return hipGetLastError();
}

int oz_narrow_digits() { return OZ_NARROW_TB; }

hipError_t oz_gemm_narrow(hipStream_t s, int64_t m, int64_t n, int64_t K, const int8_t *dA, int64_t planeA, int SA, const int8_t *dB,
                          int64_t ldb, int64_t planeB, int32_t *out) {
    if (m == 0 || n == 0) return hipSuccess;
    if (K > 8192 || K % OZ_KS != 0 || ldb % 16 != 0 || planeA < cdiv(m, 16) * 16 * K) return hipErrorInvalidValue;
    const dim3 grid((unsigned)cdiv(m, 16), (unsigned)cdiv(n, 64)), block(256);
    switch (SA) {   // the digit counts of oz_digits(N)
        case 34: hipLaunchKernelGGL((k_oz_gemm_narrow<34, OZ_NARROW_TB>), grid, block, 0, s, m, n, K, dA, planeA, dB, ldb, planeB, out); break;
        case 27: hipLaunchKernelGGL((k_oz_gemm_narrow<27, OZ_NARROW_TB>), grid, block, 0, s, m, n, K, dA, planeA, dB, ldb, planeB, out); break;
        case 19: hipLaunchKernelGGL((k_oz_gemm_narrow<19, OZ_NARROW_TB>), grid, block, 0, s, m, n, K, dA, planeA, dB, ldb, planeB, out); break;
        case 11: hipLaunchKernelGGL((k_oz_gemm_narrow<11, OZ_NARROW_TB>), grid, block, 0, s, m, n, K, dA, planeA, dB, ldb, planeB, out); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t oz_round_fp64(hipStream_t s, int64_t K, int64_t ncols, const DV &D, const int *e, int T, int8_t *dig, int64_t ld,
                         int64_t plane) {
    if (ld * ncols == 0) return hipSuccess;
    if (T < 1 || T > 8 || ld % 64 != 0 || ld < K) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_oz_round_fp64, dim3((unsigned)cdiv(ld * ncols, 256)), dim3(256), 0, s, K, ncols, D, e, T, dig, ld, plane);
    return hipGetLastError();
}

hipError_t oz_gemm(hipStream_t s, int64_t m, int64_t n, int64_t K, const int8_t *dA, int64_t lda, int64_t planeA, int SA,
                   const int8_t *dB, int64_t ldb, int64_t planeB, int SB, int U, int32_t *out) {
    if (m == 0 || n == 0) return hipSuccess;
    // int32 headroom: pairs * K * 2^12 < 2^31
    if (K > 8192 || K % OZ_KS != 0 || lda % 16 != 0 || ldb % 16 != 0 || (SA < SB ? SA : SB) > 34 || U > SA + SB - 2)
        return hipErrorInvalidValue;
    if (n <= 64) {
        const int tiles_m = (int)cdiv(m, 128);
        const dim3 grid((unsigned)(tiles_m * cdiv(n, 64)), 1, (unsigned)(U + 1));
        hipLaunchKernelGGL((k_oz_gemm<4, 2>), grid, dim3(256), 0, s, m, n, K, dA, lda, planeA, SA, dB, ldb, planeB, SB, out, tiles_m);
    } else {
        const int tiles_m = (int)cdiv(m, 128);
        const dim3 grid((unsigned)(tiles_m * cdiv(n, 128)), 1, (unsigned)(U + 1));
        hipLaunchKernelGGL((k_oz_gemm<4, 4>), grid, dim3(256), 0, s, m, n, K, dA, lda, planeA, SA, dB, ldb, planeB, SB, out, tiles_m);
    }
    return hipGetLastError();
}

int oz_row_align() { return OZ_KS; }

hipError_t oz_combine(int N, hipStream_t s, int64_t m, int64_t n, int U, const int32_t *planes, const int *ea, const int *eb,
                      const DV &C, const double *alpha, const double *beta, int beta_zero, int mode) {
    if (m * n == 0) return hipSuccess;
    OZ_DISPATCH(N, hipLaunchKernelGGL(k_oz_combine<NN>, dim3((unsigned)cdiv(m * n, 256)), dim3(256), 0, s, m, n, U, planes, ea, eb, C,
                                      mk_oz<NN>(alpha), mk_oz<NN>(beta), beta_zero, mode));
    return hipGetLastError();
}

}  // namespace mfk
