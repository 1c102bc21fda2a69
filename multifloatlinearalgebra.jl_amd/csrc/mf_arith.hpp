// Multi-double (Float64xN, N = 1..4) arithmetic for CDNA4 (gfx950) and host.
//
// An element is N binary64 limbs, most significant first, non-overlapping.
// Every network here is a fixed straight-line sequence of binary64 add, mul
// and fma, so a given input produces the same limbs on the GPU and the host.
// All translation units including this header are compiled with
// -ffp-contract=off: contracting a*b+c into an FMA silently breaks the
// error-free transforms (TwoSum / TwoProd) the networks are built from.
//
// Network provenance (reference = MultiFloatLinearAlgebra.jl, "MFLA"):
//   x3 add         src/kernels/mulacc_x3.jl:41-59 (_raw_mfadd3)
//   x3 mul         prefix src/kernels/mulacc_x3.jl:20-39 + compression tail
//                  benchmark/fused_mulacc_x3.jl:64-73 (":current")
//   x3 fused mac   src/kernels/mulacc_x3.jl:61-66 (mulacc_x3)
//   x4 add         benchmark/ordered_twosum.jl:46-81 (ordered_two_sum == two_sum)
//   x2 add/mul, x4 mul, div, inv, sqrt: MultiFloats.jl's own networks are not
//   in the reference tree; these are this project's networks (x4 mul after
//   the Hida-Li-Bailey quad-double product, long division, Heron sqrt).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define MF_HD __host__ __device__ __forceinline__

namespace mf {

template <int N> struct V { double x[N]; };

MF_HD void two_sum(double a, double b, double &s, double &e) {
    s = a + b;
    double ap = s - b;
    double bp = s - ap;
    e = (a - ap) + (b - bp);
}
MF_HD void fast_two_sum(double a, double b, double &s, double &e) {
    s = a + b;
    double bp = s - a;
    e = b - bp;
}
MF_HD void two_prod(double a, double b, double &p, double &e) {
    p = a * b;
    e = __builtin_fma(a, b, -p);
}

template <int N> MF_HD V<N> zero() { V<N> r; _Pragma("unroll") for (int i = 0; i < N; ++i) r.x[i] = 0.0; return r; }
template <int N> MF_HD V<N> from(double v) { V<N> r = zero<N>(); r.x[0] = v; return r; }
template <int N> MF_HD V<N> neg(const V<N> &a) { V<N> r; _Pragma("unroll") for (int i = 0; i < N; ++i) r.x[i] = -a.x[i]; return r; }

// ------------------------------------------------------------------ add
MF_HD V<1> add(const V<1> &x, const V<1> &y) { V<1> r; r.x[0] = x.x[0] + y.x[0]; return r; }

MF_HD V<2> add(const V<2> &x, const V<2> &y) {
    double a, b, c, d;
    two_sum(x.x[0], y.x[0], a, b);
    two_sum(x.x[1], y.x[1], c, d);
    fast_two_sum(a, c, a, c);
    b += d;
    b += c;
    fast_two_sum(a, b, a, b);
    V<2> r; r.x[0] = a; r.x[1] = b; return r;
}

MF_HD V<3> add3_raw(double x0, double x1, double x2, double y0, double y1, double y2) {
    double a, b, c, d, e, f;
    two_sum(x0, y0, a, b);
    two_sum(x1, y1, c, d);
    two_sum(x2, y2, e, f);
    fast_two_sum(a, c, a, c);
    b += f;
    two_sum(d, e, d, e);
    fast_two_sum(a, d, a, d);
    two_sum(b, c, b, c);
    c += e;
    two_sum(c, d, c, d);
    two_sum(b, c, b, c);
    fast_two_sum(a, b, a, b);
    c += d;
    fast_two_sum(b, c, b, c);
    fast_two_sum(a, b, a, b);
    fast_two_sum(b, c, b, c);
    V<3> r; r.x[0] = a; r.x[1] = b; r.x[2] = c; return r;
}
MF_HD V<3> add(const V<3> &x, const V<3> &y) { return add3_raw(x.x[0], x.x[1], x.x[2], y.x[0], y.x[1], y.x[2]); }

MF_HD V<4> add(const V<4> &x, const V<4> &y) {
    double a, b, c, d, e, f, g, h;
    two_sum(x.x[0], y.x[0], a, b);
    two_sum(x.x[1], y.x[1], c, d);
    two_sum(x.x[2], y.x[2], e, f);
    two_sum(x.x[3], y.x[3], g, h);
    fast_two_sum(a, c, a, c);
    b += h;
    two_sum(d, e, d, e);
    two_sum(f, g, f, g);
    two_sum(b, g, b, g);
    fast_two_sum(c, d, c, d);
    two_sum(e, f, e, f);
    fast_two_sum(a, c, a, c);
    fast_two_sum(d, e, d, e);
    two_sum(b, d, b, d);
    fast_two_sum(c, g, c, g);
    e += f;
    two_sum(b, c, b, c);
    two_sum(d, e, d, e);
    fast_two_sum(a, b, a, b);
    two_sum(c, d, c, d);
    e += g;
    fast_two_sum(b, c, b, c);
    two_sum(d, e, d, e);
    fast_two_sum(a, b, a, b);
    fast_two_sum(c, d, c, d);
    fast_two_sum(b, c, b, c);
    d += e;
    fast_two_sum(a, b, a, b);
    fast_two_sum(c, d, c, d);
    fast_two_sum(b, c, b, c);
    fast_two_sum(c, d, c, d);
    V<4> r; r.x[0] = a; r.x[1] = b; r.x[2] = c; r.x[3] = d; return r;
}

template <int N> MF_HD V<N> sub(const V<N> &x, const V<N> &y) { return add(x, neg(y)); }

// ------------------------------------------------------------------ mul
MF_HD V<1> mul(const V<1> &x, const V<1> &y) { V<1> r; r.x[0] = x.x[0] * y.x[0]; return r; }

MF_HD V<2> mul(const V<2> &x, const V<2> &y) {
    double p, e;
    two_prod(x.x[0], y.x[0], p, e);
    // first-order cross terms fused into the error term (two FMAs = the same
    // two multiplies and two additions, one rounding each)
    e = __builtin_fma(x.x[0], y.x[1], e);
    e = __builtin_fma(x.x[1], y.x[0], e);
    fast_two_sum(p, e, p, e);
    V<2> r; r.x[0] = p; r.x[1] = e; return r;
}

MF_HD void prefix3(const V<3> &x, const V<3> &y, double &o0, double &o1, double &o2) {
    double p00, e00, p01, e01, p10, e10;
    two_prod(x.x[0], y.x[0], p00, e00);
    two_prod(x.x[0], y.x[1], p01, e01);
    two_prod(x.x[1], y.x[0], p10, e10);
    double p02 = x.x[0] * y.x[2];
    double p11 = x.x[1] * y.x[1];
    double p20 = x.x[2] * y.x[0];
    two_sum(p01, p10, p01, p10);
    e01 += e10;
    p02 += p20;
    two_sum(e00, p01, e00, p01);
    p02 += p11;
    fast_two_sum(p00, e00, p00, e00);
    p01 += p10;
    e01 += p02;
    p01 += e01;
    o0 = p00; o1 = e00; o2 = p01;
}

MF_HD V<3> mul(const V<3> &x, const V<3> &y) {
    double p0, p1, p2;
    prefix3(x, y, p0, p1, p2);
    two_sum(p1, p2, p1, p2);
    fast_two_sum(p0, p1, p0, p1);
    fast_two_sum(p1, p2, p1, p2);
    fast_two_sum(p0, p1, p0, p1);
    V<3> r; r.x[0] = p0; r.x[1] = p1; r.x[2] = p2; return r;
}

MF_HD V<4> mul(const V<4> &x, const V<4> &y) {
    double p00, e00, p01, e01, p10, e10, p02, e02, p11, e11, p20, e20;
    two_prod(x.x[0], y.x[0], p00, e00);
    two_prod(x.x[0], y.x[1], p01, e01);
    two_prod(x.x[1], y.x[0], p10, e10);
    two_prod(x.x[0], y.x[2], p02, e02);
    two_prod(x.x[1], y.x[1], p11, e11);
    two_prod(x.x[2], y.x[0], p20, e20);
    double s1, t1, t2, p0;
    two_sum(p01, p10, s1, t1);
    two_sum(e00, s1, s1, t2);
    fast_two_sum(p00, s1, p0, s1);
    double a, b, c, d, f, g, h, i, j;
    two_sum(p02, p20, a, b);
    two_sum(a, p11, a, c);
    two_sum(e01, e10, d, f);
    two_sum(a, d, a, g);
    two_sum(t1, t2, h, i);
    two_sum(a, h, a, j);
    // third-order terms: the rounded errors of the lower orders summed, then
    // the four third-order products fused in (FMA: one multiply and one
    // addition each)
    double z = e02 + e20;
    z += e11;
    double w = b + c;
    z += w;
    w = f + g;
    z += w;
    w = i + j;
    z += w;
    z = __builtin_fma(x.x[0], y.x[3], z);
    z = __builtin_fma(x.x[3], y.x[0], z);
    z = __builtin_fma(x.x[1], y.x[2], z);
    z = __builtin_fma(x.x[2], y.x[1], z);

    two_sum(s1, a, s1, a);
    two_sum(a, z, a, z);
    fast_two_sum(p0, s1, p0, s1);
    fast_two_sum(s1, a, s1, a);
    fast_two_sum(a, z, a, z);
    fast_two_sum(p0, s1, p0, s1);
    fast_two_sum(s1, a, s1, a);
    fast_two_sum(p0, s1, p0, s1);
    V<4> r; r.x[0] = p0; r.x[1] = s1; r.x[2] = a; r.x[3] = z; return r;
}

// ------------------------------------------------- multiply-accumulate
template <int N> MF_HD V<N> mac(const V<N> &acc, const V<N> &x, const V<N> &y) { return add(acc, mul(x, y)); }
template <int N> MF_HD V<N> msub(const V<N> &acc, const V<N> &x, const V<N> &y) { return add(acc, neg(mul(x, y))); }

MF_HD V<3> mac_fused3(const V<3> &acc, const V<3> &x, const V<3> &y) {
    double p0, p1, p2;
    prefix3(x, y, p0, p1, p2);
    two_sum(p1, p2, p1, p2);
    return add3_raw(acc.x[0], acc.x[1], acc.x[2], p0, p1, p2);
}

// Network selector for GEMM-family kernels: the fused x3 accumulation is used
// only where MFLA's GEMM uses it (four-row vector blocks of the x3 route).
enum Net { NET_STD = 0, NET_FUSED3 = 1 };
template <int N, int NET> struct Mac {
    static MF_HD V<N> apply(const V<N> &acc, const V<N> &x, const V<N> &y) { return mac(acc, x, y); }
};
template <> struct Mac<3, NET_FUSED3> {
    static MF_HD V<3> apply(const V<3> &acc, const V<3> &x, const V<3> &y) { return mac_fused3(acc, x, y); }
};

// --------------------------------------------------------- comparisons
// Lexicographic over the limbs: the first limb where x < y or x > y decides
// (a NaN limb decides nothing and passes to the next). Written without early
// returns: the per-limb early-return form, used as the condition of a
// running argmax update in a loop, was compiled (ROCm 7.2, gfx950, N >= 2)
// into code that stopped updating the leading limb of the running value.
template <int N> MF_HD bool lt(const V<N> &x, const V<N> &y) {
    bool less = false, open = true;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const bool l = x.x[i] < y.x[i], g = x.x[i] > y.x[i];
        less = less || (open && l);
        open = open && !l && !g;
    }
    return less;
}
template <int N> MF_HD bool gt(const V<N> &x, const V<N> &y) { return lt(y, x); }
template <int N> MF_HD bool eq(const V<N> &x, const V<N> &y) {
    bool same = true;
#pragma unroll
    for (int i = 0; i < N; ++i) same = same && x.x[i] == y.x[i];
    return same;
}
template <int N> MF_HD bool ge(const V<N> &x, const V<N> &y) { return gt(x, y) || eq(x, y); }
template <int N> MF_HD bool is_zero(const V<N> &x) { for (int i = 0; i < N; ++i) if (x.x[i] != 0.0) return false; return true; }
template <int N> MF_HD bool is_finite(const V<N> &x) { for (int i = 0; i < N; ++i) if (!__builtin_isfinite(x.x[i])) return false; return true; }
template <int N> MF_HD bool is_one(const V<N> &x) { if (x.x[0] != 1.0) return false; for (int i = 1; i < N; ++i) if (x.x[i] != 0.0) return false; return true; }
template <int N> MF_HD V<N> vabs(const V<N> &x) { return __builtin_signbit(x.x[0]) ? neg(x) : x; }
template <int N> MF_HD V<N> vmax(const V<N> &x, const V<N> &y) { return lt(y, x) ? x : y; }
template <int N> MF_HD V<N> vmin(const V<N> &x, const V<N> &y) { return lt(y, x) ? y : x; }

// ----------------------------------------------------- division / sqrt
template <int N> MF_HD V<N> renorm_np1(double *t) {
    for (int i = N; i > 0; --i) two_sum(t[i - 1], t[i], t[i - 1], t[i]);
    for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i < N; ++i) fast_two_sum(t[i], t[i + 1], t[i], t[i + 1]);
    V<N> r;
    for (int i = 0; i < N; ++i) r.x[i] = t[i];
    if (N > 1) r.x[N - 1] += t[N];
    return r;
}
// a (N limbs) * d (binary64) to N-limb accuracy: TwoProd of every limb but
// the last, the partial products gathered by magnitude level (TwoSum where a
// level has two terms), then FastTwoSum sweeps.
template <int N> MF_HD V<N> mul_dbl(const V<N> &a, double d);
template <> MF_HD V<1> mul_dbl<1>(const V<1> &a, double d) { V<1> r; r.x[0] = a.x[0] * d; return r; }
template <> MF_HD V<2> mul_dbl<2>(const V<2> &a, double d) {
    double p0, e0;
    two_prod(a.x[0], d, p0, e0);
    double s1 = a.x[1] * d;
    s1 += e0;
    fast_two_sum(p0, s1, p0, s1);
    V<2> r; r.x[0] = p0; r.x[1] = s1; return r;
}
template <> MF_HD V<3> mul_dbl<3>(const V<3> &a, double d) {
    double p0, e0, p1, e1, s1, t1;
    two_prod(a.x[0], d, p0, e0);
    two_prod(a.x[1], d, p1, e1);
    double s2 = a.x[2] * d;
    two_sum(p1, e0, s1, t1);
    s2 += e1;
    s2 += t1;
    fast_two_sum(p0, s1, p0, s1);
    fast_two_sum(s1, s2, s1, s2);
    fast_two_sum(p0, s1, p0, s1);
    V<3> r; r.x[0] = p0; r.x[1] = s1; r.x[2] = s2; return r;
}
template <> MF_HD V<4> mul_dbl<4>(const V<4> &a, double d) {
    double p0, e0, p1, e1, p2, e2, s1, t1, s2, t2, t3;
    two_prod(a.x[0], d, p0, e0);
    two_prod(a.x[1], d, p1, e1);
    two_prod(a.x[2], d, p2, e2);
    double s3 = a.x[3] * d;
    two_sum(p1, e0, s1, t1);
    two_sum(p2, e1, s2, t2);
    two_sum(s2, t1, s2, t3);
    s3 += e2;
    t2 += t3;
    s3 += t2;
    fast_two_sum(p0, s1, p0, s1);
    fast_two_sum(s1, s2, s1, s2);
    fast_two_sum(s2, s3, s2, s3);
    fast_two_sum(p0, s1, p0, s1);
    fast_two_sum(s1, s2, s1, s2);
    fast_two_sum(p0, s1, p0, s1);
    V<4> r; r.x[0] = p0; r.x[1] = s1; r.x[2] = s2; r.x[3] = s3; return r;
}

// Long division: N+1 partial quotients q_i = r0 / y0, the remainder updated
// as r - y*q_i with the N-limb x binary64 product, then an error-free
// renormalisation of the quotient digits.
template <int N> MF_HD V<N> div(const V<N> &x, const V<N> &y) {
    if constexpr (N == 1) {
        V<1> r; r.x[0] = x.x[0] / y.x[0]; return r;
    } else {
        double q[N + 1];
        V<N> r = x;
        for (int i = 0; i <= N; ++i) {
            q[i] = r.x[0] / y.x[0];
            if (i < N) r = sub(r, mul_dbl<N>(y, q[i]));
        }
        return renorm_np1<N>(q);
    }
}
template <int N> MF_HD V<N> inv(const V<N> &y) { return div(from<N>(1.0), y); }
template <int N> MF_HD V<N> sqrt(const V<N> &x) {
    if (!(x.x[0] > 0.0)) {
        if (x.x[0] == 0.0) return zero<N>();
        return from<N>(__builtin_nan(""));
    }
    V<N> s = from<N>(::sqrt(x.x[0]));
    constexpr int iters = N == 1 ? 0 : (N == 2 ? 1 : 2);
    for (int it = 0; it < iters; ++it) {
        s = add(s, div(x, s));
        for (int i = 0; i < N; ++i) s.x[i] *= 0.5;
    }
    return s;
}

}  // namespace mf
