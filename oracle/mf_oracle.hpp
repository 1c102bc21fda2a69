// TEST INFRASTRUCTURE ONLY. This header is part of the CPU oracle: a plain C++
// restatement of MultiFloatLinearAlgebra.jl (MFLA) and of the multi-double
// arithmetic networks it runs on. Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load the oracle, and only as the checker.
// Nothing in the product library includes or links this file.
//
// Compile every translation unit that includes this file with
// -ffp-contract=off and without fast-math: contraction of a*b+c into an FMA
// silently destroys the error-free transforms below.
//
// Element representation: MF<N> holds N binary64 limbs, most significant first.
//
// Parity status of each network (what pins it):
//   two_sum / fast_two_sum / two_prod  -- textbook EFTs; fast_two_sum literal
//       at benchmark/ordered_twosum.jl:10-15; two_prod uses an FMA error term
//       (docs/MULACC_X3_PROOF_CONTRACT.md:37). Pinned by exactness invariants.
//   x3 add   -- src/kernels/mulacc_x3.jl:41-59 (_raw_mfadd3). Pinned by the
//       8 proof vectors of test/mulacc_x3_proof_vectors.jl.
//   x3 mul   -- product prefix src/kernels/mulacc_x3.jl:20-39 plus the
//       ":current" compression tail benchmark/fused_mulacc_x3.jl:64-73.
//       Pinned by the same proof vectors (standard network acc + x*y).
//   x3 fused mulacc -- src/kernels/mulacc_x3.jl:61-66. Pinned by the vectors.
//   x4 add   -- benchmark/ordered_twosum.jl:46-81 with ordered_two_sum
//       replaced by two_sum (both return the exact (sum, error) pair, see
//       benchmark/ARITHMETIC_PROFILE.md "magnitude-ordered FastTwoSum").
//       Topology pinned; no committed x4 vectors exist: pinned by mpmath
//       exactness/accuracy fixtures generated in tests/golden/.
//   x2 add, x2 mul, x4 mul, div, inv, sqrt -- MultiFloats.jl 3.2.6 networks are
//       not in the reference tree: PARITY UNPINNED against MultiFloats bits.
//       The networks below are this project's own (x2 add follows the same
//       accumulation-network shape as the x3 add; x4 mul follows the
//       Hida-Li-Bailey quad-double product structure; division is long
//       division; sqrt is Heron iteration). They are pinned by accuracy
//       against mpmath at 512 bits and by closed-form identities.
#pragma once
#include <cmath>
#include <cstdint>

namespace orc {

static inline void two_sum(double a, double b, double &s, double &e) {
    s = a + b;
    double ap = s - b;
    double bp = s - ap;
    e = (a - ap) + (b - bp);
}

static inline void fast_two_sum(double a, double b, double &s, double &e) {
    s = a + b;
    double bp = s - a;
    e = b - bp;
}

static inline void two_prod(double a, double b, double &p, double &e) {
    p = a * b;
    e = std::fma(a, b, -p);
}

template <int N> struct MF { double x[N]; };

template <int N> static inline MF<N> mf_zero() { MF<N> r; for (int i = 0; i < N; ++i) r.x[i] = 0.0; return r; }
template <int N> static inline MF<N> mf_from(double v) { MF<N> r = mf_zero<N>(); r.x[0] = v; return r; }
template <int N> static inline MF<N> mf_neg(const MF<N> &a) { MF<N> r; for (int i = 0; i < N; ++i) r.x[i] = -a.x[i]; return r; }

// ---------------------------------------------------------------- addition
static inline MF<1> add(const MF<1> &x, const MF<1> &y) { MF<1> r; r.x[0] = x.x[0] + y.x[0]; return r; }

static inline MF<2> add(const MF<2> &x, const MF<2> &y) {
    double a, b, c, d;
    two_sum(x.x[0], y.x[0], a, b);
    two_sum(x.x[1], y.x[1], c, d);
    fast_two_sum(a, c, a, c);
    b += d;
    b += c;
    fast_two_sum(a, b, a, b);
    MF<2> r; r.x[0] = a; r.x[1] = b; return r;
}

// src/kernels/mulacc_x3.jl:41-59
static inline MF<3> raw_add3(double x0, double x1, double x2, double y0, double y1, double y2) {
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
    MF<3> r; r.x[0] = a; r.x[1] = b; r.x[2] = c; return r;
}
static inline MF<3> add(const MF<3> &x, const MF<3> &y) { return raw_add3(x.x[0], x.x[1], x.x[2], y.x[0], y.x[1], y.x[2]); }

// benchmark/ordered_twosum.jl:46-81 (ordered_two_sum -> two_sum)
static inline MF<4> add(const MF<4> &x, const MF<4> &y) {
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
    MF<4> r; r.x[0] = a; r.x[1] = b; r.x[2] = c; r.x[3] = d; return r;
}

template <int N> static inline MF<N> sub(const MF<N> &x, const MF<N> &y) { return add(x, mf_neg(y)); }

// ---------------------------------------------------------- multiplication
static inline MF<1> mul(const MF<1> &x, const MF<1> &y) { MF<1> r; r.x[0] = x.x[0] * y.x[0]; return r; }

static inline MF<2> mul(const MF<2> &x, const MF<2> &y) {
    double p, e;
    two_prod(x.x[0], y.x[0], p, e);
    // first-order cross terms fused into the error term (two FMAs = the same
    // two multiplies and two additions, one rounding each)
    e = __builtin_fma(x.x[0], y.x[1], e);
    e = __builtin_fma(x.x[1], y.x[0], e);
    fast_two_sum(p, e, p, e);
    MF<2> r; r.x[0] = p; r.x[1] = e; return r;
}

// src/kernels/mulacc_x3.jl:20-39
static inline void product_prefix3(const MF<3> &x, const MF<3> &y, double &o0, double &o1, double &o2) {
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

// prefix + ":current" tail, benchmark/fused_mulacc_x3.jl:64-73
static inline MF<3> mul(const MF<3> &x, const MF<3> &y) {
    double p0, p1, p2;
    product_prefix3(x, y, p0, p1, p2);
    two_sum(p1, p2, p1, p2);
    fast_two_sum(p0, p1, p0, p1);
    fast_two_sum(p1, p2, p1, p2);
    fast_two_sum(p0, p1, p0, p1);
    MF<3> r; r.x[0] = p0; r.x[1] = p1; r.x[2] = p2; return r;
}

// Quad-double product (this project's network; structure after
// Hida, Li & Bailey 2001): exact products for the three leading diagonals,
// rounded products for the fourth, an error-free level-by-level accumulation,
// then a fast_two_sum renormalisation sweep. Symmetric in x and y.
static inline MF<4> mul(const MF<4> &x, const MF<4> &y) {
    double p00, e00, p01, e01, p10, e10, p02, e02, p11, e11, p20, e20;
    two_prod(x.x[0], y.x[0], p00, e00);
    two_prod(x.x[0], y.x[1], p01, e01);
    two_prod(x.x[1], y.x[0], p10, e10);
    two_prod(x.x[0], y.x[2], p02, e02);
    two_prod(x.x[1], y.x[1], p11, e11);
    two_prod(x.x[2], y.x[0], p20, e20);
    // first-order diagonal
    double s1, t1, t2;
    two_sum(p01, p10, s1, t1);
    two_sum(e00, s1, s1, t2);
    double p0;
    fast_two_sum(p00, s1, p0, s1);
    // second-order diagonal
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
    MF<4> r; r.x[0] = p0; r.x[1] = s1; r.x[2] = a; r.x[3] = z; return r;
}

// ------------------------------------------------------ multiply-accumulate
// Standard network acc + x*y (MultiFloats `acc + x * y`).
template <int N> static inline MF<N> mac(const MF<N> &acc, const MF<N> &x, const MF<N> &y) { return add(acc, mul(x, y)); }
// acc - x*y  (MultiFloats `acc - x*y` = acc + (-(x*y)))
template <int N> static inline MF<N> msub(const MF<N> &acc, const MF<N> &x, const MF<N> &y) { return add(acc, mf_neg(mul(x, y))); }

// Fused x3 multiply-accumulate, src/kernels/mulacc_x3.jl:61-66 (the "mid" cut).
static inline MF<3> mulacc_x3(const MF<3> &acc, const MF<3> &x, const MF<3> &y) {
    double p0, p1, p2;
    product_prefix3(x, y, p0, p1, p2);
    two_sum(p1, p2, p1, p2);
    return raw_add3(acc.x[0], acc.x[1], acc.x[2], p0, p1, p2);
}

// ---------------------------------------------------------- comparisons
template <int N> static inline bool lt(const MF<N> &x, const MF<N> &y) {
    for (int i = 0; i < N; ++i) {
        if (x.x[i] < y.x[i]) return true;
        if (x.x[i] > y.x[i]) return false;
    }
    return false;
}
template <int N> static inline bool gt(const MF<N> &x, const MF<N> &y) { return lt(y, x); }
template <int N> static inline bool eq(const MF<N> &x, const MF<N> &y) {
    for (int i = 0; i < N; ++i) if (!(x.x[i] == y.x[i])) return false;
    return true;
}
template <int N> static inline bool ge(const MF<N> &x, const MF<N> &y) { return gt(x, y) || eq(x, y); }
template <int N> static inline bool iszero(const MF<N> &x) { for (int i = 0; i < N; ++i) if (x.x[i] != 0.0) return false; return true; }
template <int N> static inline bool isfin(const MF<N> &x) { for (int i = 0; i < N; ++i) if (!std::isfinite(x.x[i])) return false; return true; }
template <int N> static inline bool is_one(const MF<N> &x) { if (x.x[0] != 1.0) return false; for (int i = 1; i < N; ++i) if (x.x[i] != 0.0) return false; return true; }
template <int N> static inline MF<N> mf_abs(const MF<N> &x) { return std::signbit(x.x[0]) ? mf_neg(x) : x; }
// Julia: max(x, y) = isless(y, x) ? x : y ; min(x, y) = isless(y, x) ? y : x
template <int N> static inline MF<N> mf_max(const MF<N> &x, const MF<N> &y) { return lt(y, x) ? x : y; }
template <int N> static inline MF<N> mf_min(const MF<N> &x, const MF<N> &y) { return lt(y, x) ? y : x; }

// ---------------------------------------------------------- division / sqrt
// Long division: N+1 partial quotients q_i = r0 / y0 with the remainder
// r - y*q_i updated in N-limb arithmetic (y*q_i by the N-limb x binary64
// product below), then an error-free renormalisation.
template <int N> static inline MF<N> renorm_from(const double *q) {
    // q has N+1 terms of decreasing magnitude: bottom-up two_sum, then two
    // top-down fast_two_sum sweeps; the last term is folded into limb N-1.
    double t[N + 1];
    for (int i = 0; i <= N; ++i) t[i] = q[i];
    for (int i = N; i > 0; --i) two_sum(t[i - 1], t[i], t[i - 1], t[i]);
    for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i < N; ++i) fast_two_sum(t[i], t[i + 1], t[i], t[i + 1]);
    MF<N> r;
    for (int i = 0; i < N; ++i) r.x[i] = t[i];
    if (N > 1) r.x[N - 1] += t[N];
    return r;
}
// y (N limbs) times a binary64 q to N-limb accuracy: the exact products of
// the leading N-1 limbs (TwoProd) and the plain product of the last, summed
// by magnitude level -- level 1: p1 + e0 (TwoSum), level 2: p2 + e1 + the
// level-1 error (TwoSum, TwoSum), level 3: p3 + e2 + the level-2 errors
// (plain) -- then FastTwoSum sweeps (p0,s1),(s1,s2),(s2,s3),(p0,s1),(s1,s2),
// (p0,s1).
template <int N> static inline MF<N> mul_by_double(const MF<N> &y, double q) {
    MF<N> r;
    if (N == 1) { r.x[0] = y.x[0] * q; return r; }
    double p0, e0;
    two_prod(y.x[0], q, p0, e0);
    if (N == 2) {
        double s1 = y.x[1] * q;
        s1 = s1 + e0;
        fast_two_sum(p0, s1, p0, s1);
        r.x[0] = p0; r.x[1] = s1;
        return r;
    }
    double p1, e1, s1, t1;
    two_prod(y.x[1], q, p1, e1);
    if (N == 3) {
        double s2 = y.x[2] * q;
        two_sum(p1, e0, s1, t1);
        s2 = s2 + e1;
        s2 = s2 + t1;
        fast_two_sum(p0, s1, p0, s1);
        fast_two_sum(s1, s2, s1, s2);
        fast_two_sum(p0, s1, p0, s1);
        r.x[0] = p0; r.x[1] = s1; r.x[2] = s2;
        return r;
    }
    double p2, e2, s2, t2, t3;
    two_prod(y.x[2], q, p2, e2);
    double s3 = y.x[3] * q;
    two_sum(p1, e0, s1, t1);
    two_sum(p2, e1, s2, t2);
    two_sum(s2, t1, s2, t3);
    s3 = s3 + e2;
    t2 = t2 + t3;
    s3 = s3 + t2;
    fast_two_sum(p0, s1, p0, s1);
    fast_two_sum(s1, s2, s1, s2);
    fast_two_sum(s2, s3, s2, s3);
    fast_two_sum(p0, s1, p0, s1);
    fast_two_sum(s1, s2, s1, s2);
    fast_two_sum(p0, s1, p0, s1);
    r.x[0] = p0; r.x[1] = s1; r.x[2] = s2; r.x[3] = s3;
    return r;
}
template <int N> static inline MF<N> mf_div(const MF<N> &x, const MF<N> &y) {
    if (N == 1) { MF<N> r; r.x[0] = x.x[0] / y.x[0]; return r; }
    double q[N + 1];
    MF<N> r = x;
    for (int i = 0; i <= N; ++i) {
        q[i] = r.x[0] / y.x[0];
        if (i < N) r = sub(r, mul_by_double(y, q[i]));
    }
    return renorm_from<N>(q);
}
template <int N> static inline MF<N> mf_inv(const MF<N> &y) { return mf_div(mf_from<N>(1.0), y); }
template <int N> static inline MF<N> mf_sqrt(const MF<N> &x) {
    if (!(x.x[0] > 0.0)) {
        if (x.x[0] == 0.0) return mf_zero<N>();
        return mf_from<N>(std::nan(""));
    }
    MF<N> s = mf_from<N>(std::sqrt(x.x[0]));
    const int iters = N == 1 ? 0 : (N == 2 ? 1 : 2);
    for (int it = 0; it < iters; ++it) {
        s = add(s, mf_div(x, s));
        for (int i = 0; i < N; ++i) s.x[i] *= 0.5;
    }
    return s;
}

}  // namespace orc
