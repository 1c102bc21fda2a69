// TEST INFRASTRUCTURE ONLY -- exact reference arithmetic for the oracle.
// A Kulisch-style fixed-point accumulator covering 2^-2200 .. 2^2100 adds
// binary64 values and exact products of binary64 pairs with no rounding, so
// sums of products of multi-double limbs are evaluated exactly. It stands in
// for MFLA's BigFloat-512 checks (test/runtests.jl:10-23) at sizes where
// mpmath is too slow. tests/test_oracle_exact.py checks it against exact
// rational arithmetic (fractions.Fraction): wide exponents, full
// cancellation, both signs, transposed addressing, range edges.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdlib>

namespace {

constexpr int kBase = 2200;          // bit position 0 of chunk 0 is 2^-kBase
constexpr int kChunks = 140;         // 140 * 32 = 4480 bits

struct Acc {
    int64_t c[kChunks];
    void clear() { std::memset(c, 0, sizeof(c)); }
    // add sign * mant * 2^exp (mant < 2^127)
    void add_mag(unsigned __int128 mant, int exp, int sign) {
        if (mant == 0) return;
        int pos = exp + kBase;
        if (pos < 0) {  // below range: drop bits (not reachable for binary64 products)
            mant >>= (-pos); pos = 0;
            if (mant == 0) return;
        }
        int idx = pos >> 5, sh = pos & 31;
        for (int i = 0; i < 4 && mant != 0; ++i) {
            uint64_t part = (uint64_t)(mant & 0xffffffffu);
            mant >>= 32;
            uint64_t v = part << sh;
            int64_t lo = (int64_t)(v & 0xffffffffu), hi = (int64_t)(v >> 32);
            if (sign > 0) { c[idx + i] += lo; c[idx + i + 1] += hi; }
            else { c[idx + i] -= lo; c[idx + i + 1] -= hi; }
        }
    }
    // |x| = mant * 2^exp from the binary64 fields (subnormals: exponent
    // field 0 means 2^-1074 units); false for zero, Inf and NaN.
    static bool split(double x, uint64_t &mant, int &exp) {
        uint64_t b;
        std::memcpy(&b, &x, sizeof(b));
        const int ef = (int)((b >> 52) & 0x7ff);
        mant = b & ((1ull << 52) - 1);
        if (ef == 0x7ff) return false;
        if (ef == 0) { exp = -1074; return mant != 0; }
        mant |= 1ull << 52;
        exp = ef - 1075;
        return true;
    }
    void add_double(double x) {
        uint64_t m; int e;
        if (!split(x, m, e)) return;
        add_mag((unsigned __int128)m, e, std::signbit(x) ? -1 : 1);
    }
    void add_prod(double a, double b, int sign) {
        uint64_t ma, mb; int ea, eb;
        if (!split(a, ma, ea) || !split(b, mb, eb)) return;
        const int s = (std::signbit(a) != std::signbit(b)) ? -1 : 1;
        add_mag((unsigned __int128)ma * mb, ea + eb, s * sign);
    }
    // normalise carries; returns sign of the value (-1, 0, 1) and magnitude as double
    double to_double() {
        int64_t t[kChunks];
        std::memcpy(t, c, sizeof(t));
        for (int i = 0; i < kChunks - 1; ++i) {
            int64_t carry = t[i] >> 32;  // arithmetic shift
            t[i] -= carry * (int64_t)4294967296LL;
            t[i + 1] += carry;
        }
        int sign = t[kChunks - 1] < 0 ? -1 : 1;
        if (sign < 0) {
            // negate every digit, then renormalise: the top digit becomes positive
            for (int i = 0; i < kChunks; ++i) t[i] = -t[i];
            for (int i = 0; i < kChunks - 1; ++i) {
                int64_t carry = t[i] >> 32;
                t[i] -= carry * (int64_t)4294967296LL;
                t[i + 1] += carry;
            }
        }
        int top = kChunks - 1;
        while (top >= 0 && t[top] == 0) --top;
        if (top < 0) return 0.0;
        double r = 0.0;
        for (int i = top; i >= 0 && i >= top - 3; --i) r += std::ldexp((double)t[i], 32 * i - kBase);
        return sign * r;
    }
};

}  // namespace

extern "C" {

// For each sampled (row, col): value = sc*C[row,col] + sab * sum_k A[row,k]*B[k,col]
// evaluated exactly over all limb products, returned rounded to binary64,
// plus scale = |C| + sum_k |A||B| (binary64, approximate).
// Planar layout: element (i, j), limb l at p[l*ls + j*ld + i]. Samples are
// independent: OpenMP threads take them in chunks, one accumulator each.
int orc_exact_axpb(int N, int64_t k, const double *A, int64_t lda, int64_t lsa, int atrans, const double *B, int64_t ldb,
                   int64_t lsb, const double *C, int64_t ldc, int64_t lsc, int sab, int sc, const int64_t *rows,
                   const int64_t *cols, int64_t count, double *out_val, double *out_scale) {
    int rc = 0;
#pragma omp parallel
    {
    Acc *acc = (Acc *)std::malloc(sizeof(Acc));
    if (!acc) {
#pragma omp atomic write
        rc = -1;
    }
#pragma omp for schedule(dynamic, 16)
    for (int64_t t = 0; t < count; ++t) {
        if (!acc) continue;
        int64_t i = rows[t], j = cols[t];
        acc->clear();
        double scale = 0.0;
        if (C && sc != 0) {
            double cabs = 0.0;
            for (int l = 0; l < N; ++l) { double v = C[l * lsc + j * ldc + i]; acc->add_double(sc > 0 ? v : -v); cabs += v; }
            scale += std::fabs(cabs);
        }
        for (int64_t kk = 0; kk < k; ++kk) {
            double av = 0.0, bv = 0.0;
            for (int la = 0; la < N; ++la) {
                double a = atrans ? A[la * lsa + i * lda + kk] : A[la * lsa + kk * lda + i];
                av += a;
                for (int lb = 0; lb < N; ++lb) acc->add_prod(a, B[lb * lsb + j * ldb + kk], sab);
            }
            for (int lb = 0; lb < N; ++lb) bv += B[lb * lsb + j * ldb + kk];
            scale += std::fabs(av) * std::fabs(bv);
        }
        out_val[t] = acc->to_double();
        out_scale[t] = scale;
    }
    std::free(acc);
    }
    return rc;
}

// Exact value of sum_i x_i (count binary64 values) minus sum_j y_j, rounded.
double orc_exact_sum_diff(const double *x, int64_t nx, const double *y, int64_t ny) {
    Acc *acc = (Acc *)std::malloc(sizeof(Acc));
    acc->clear();
    for (int64_t i = 0; i < nx; ++i) acc->add_double(x[i]);
    for (int64_t i = 0; i < ny; ++i) acc->add_double(-y[i]);
    double r = acc->to_double();
    std::free(acc);
    return r;
}

}  // extern "C"
