"""The oracle's exact long accumulator (``exact_axpb``, ``exact_sum_diff``),
pinned against exact rational arithmetic on the CPU.

Every tolerance check in the GPU suite that is not bitwise (blocked TRSM,
refinement, batched systolic solves, Ozaki, QR, TRMM, the finite-difference
gap) measures its error with this accumulator, so it stands in for MFLA's
BigFloat-512 ``max_relative_error`` (test/runtests.jl:10-23). Here its value
is compared with ``fractions.Fraction`` sums of the same limb products:

* x2 / x3 / x4 operands with exponents spread over 2^-400 .. 2^400;
* full cancellation: C = fl(A B) from the oracle's gemm, so sc*C + sab*A B is
  only the rounding error of that product (both signs of sab and sc);
* the transposed-A addressing (atrans) and a missing C;
* binary64 sums with cancellation across 2^-1000 .. 2^1000.

The comparison itself is checked: a sign-flipped value, a value that drops
the lowest limb of each operand, and zero must all fail it.
"""
from fractions import Fraction

import numpy as np
import pytest

from mfutil import rand_mf

# to_double sums the top four 32-bit digits of the normalised accumulator in
# binary64: a few ulps, far below any difference the checks must detect.
REL = 2.0**-48


def close(val, exact):
    if exact == 0:
        return val == 0.0
    return abs(Fraction(float(val)) - exact) <= Fraction(REL) * abs(exact)


def frac_limbs(v):
    return sum((Fraction(float(x)) for x in v), Fraction(0))


def frac_axpb(A, B, C, sab, sc, i, j, atrans, limbs=None):
    """sc*C[i,j] + sab*sum_k op(A)[i,k] B[k,j] exactly (planar (N, cols, rows))."""
    N = A.shape[0]
    L = N if limbs is None else limbs
    k = B.shape[2]
    tot = Fraction(0)
    if C is not None and sc != 0:
        tot += sc * frac_limbs(C[:L, j, i])
    for kk in range(k):
        a = frac_limbs(A[:L, i, kk] if atrans else A[:L, kk, i])
        b = frac_limbs(B[:L, j, kk])
        tot += sab * a * b
    return tot


def test_close_rejects_planted_mistakes():
    x = Fraction(3, 7) * Fraction(2) ** -190
    assert close(float(x), x)
    assert not close(-float(x), x)
    assert not close(0.0, x)
    assert not close(float(x) * (1 + 2.0**-40), x)
    assert close(0.0, Fraction(0)) and not close(2.0**-1000, Fraction(0))


@pytest.mark.parametrize("N", [2, 3, 4])
@pytest.mark.parametrize("atrans", [False, True])
def test_exact_axpb_matches_fractions(oracle, N, atrans):
    rng = np.random.default_rng(4100 + 10 * N + int(atrans))
    m, k, n = 6, 13, 5
    A = rand_mf(rng, N, k, m, spread=400) if atrans else rand_mf(rng, N, m, k, spread=400)
    B = rand_mf(rng, N, k, n, spread=400)
    C = rand_mf(rng, N, m, n, spread=400)
    rows = np.repeat(np.arange(m), n)
    cols = np.tile(np.arange(n), m)
    for sab, sc in ((1, 1), (-1, 1), (1, -1), (-1, -1), (1, 0)):
        val, scale = oracle.exact_axpb(A, B, C, sab=sab, sc=sc, rows=rows, cols=cols, atrans=atrans)
        for t, (i, j) in enumerate(zip(rows, cols)):
            ex = frac_axpb(A, B, C, sab, sc, i, j, atrans)
            assert close(val[t], ex), (sab, sc, i, j, val[t], float(ex))
            # the magnitude it reports is |C| + sum |A||B| (binary64 estimate)
            mag = abs(frac_limbs(C[:, j, i])) * abs(sc) + sum(
                abs(frac_limbs(A[:, i, kk] if atrans else A[:, kk, i])) * abs(frac_limbs(B[:, j, kk])) for kk in range(k))
            assert abs(Fraction(float(scale[t])) - mag) <= Fraction(2.0**-40) * mag
    val0, _ = oracle.exact_axpb(A, B, None, sab=1, sc=1, rows=rows, cols=cols, atrans=atrans)
    for t, (i, j) in enumerate(zip(rows, cols)):
        assert close(val0[t], frac_axpb(A, B, None, 1, 0, i, j, atrans))


@pytest.mark.parametrize("N", [2, 3, 4])
def test_exact_axpb_full_cancellation(oracle, N):
    """C = fl(A B): the exact value A B - C is the product's rounding error,
    some 2^-(52N) below the terms; a sign, carry or dropped-limb mistake in
    the accumulator changes it completely."""
    rng = np.random.default_rng(4200 + N)
    m, k, n = 7, 21, 4
    A = rand_mf(rng, N, m, k, spread=200)
    B = rand_mf(rng, N, k, n, spread=200)
    C = oracle.gemm(np.zeros((N, n, m)), A, B)
    rows = np.repeat(np.arange(m), n)
    cols = np.tile(np.arange(n), m)
    nonzero = 0
    for sab, sc in ((1, -1), (-1, 1)):
        val, scale = oracle.exact_axpb(A, B, C, sab=sab, sc=sc, rows=rows, cols=cols)
        for t, (i, j) in enumerate(zip(rows, cols)):
            ex = frac_axpb(A, B, C, sab, sc, i, j, False)
            assert close(val[t], ex), (N, sab, sc, i, j, val[t], float(ex))
            if ex != 0:
                nonzero += 1
                # planted mistakes: each must fail the same check
                assert not close(-val[t], ex)
                assert not close(0.0, ex)
                assert not close(float(frac_axpb(A, B, C, sab, sc, i, j, False, limbs=N - 1)), ex)
            # the cancellation really happened: |A B - C| << sum |A||B|
            assert abs(val[t]) <= 2.0 ** (-52 * N + 12) * scale[t]
    assert nonzero >= len(rows)   # the rounding errors are generically non-zero


def test_exact_axpb_single_terms_at_range_edges(oracle):
    """One product or one C value at a time, at exponents near the ends of
    binary64's range and at every 32-bit digit offset: the digit placement
    and the final rounding of the accumulator."""
    N = 1
    for e in list(range(-1070, -1000, 7)) + list(range(-40, 41, 3)) + list(range(960, 1020, 7)):
        for mant in (1.0, 1.5, 1.9999999999999998, 1.2345678901234567):
            x = mant * 2.0**e
            A = np.array([[[x]]])
            B = np.array([[[1.0]]]) if e > 900 or e < -1000 else np.array([[[mant]]])
            val, _ = oracle.exact_axpb(A, B, None, sab=-1, sc=1, rows=np.array([0]), cols=np.array([0]))
            assert close(val[0], -Fraction(x) * Fraction(float(B[0, 0, 0])))
            Cv = np.array([[[x]]])
            val, _ = oracle.exact_axpb(np.zeros((1, 1, 1)), np.zeros((1, 1, 1)), Cv, sab=1, sc=-1,
                                       rows=np.array([0]), cols=np.array([0]))
            assert close(val[0], -Fraction(x))


def test_exact_sum_diff_matches_fractions(oracle):
    rng = np.random.default_rng(4300)
    for trial in range(40):
        nx, ny = rng.integers(1, 60, 2)
        x = rng.standard_normal(nx) * 2.0 ** rng.integers(-1000, 1000, nx)
        y = rng.standard_normal(ny) * 2.0 ** rng.integers(-1000, 1000, ny)
        if trial % 2:                     # cancellation: y repeats x plus one tiny term
            y = np.concatenate([x[rng.permutation(nx)], [2.0**-900 * rng.standard_normal()]])
        ex = sum(map(Fraction, map(float, x)), Fraction(0)) - sum(map(Fraction, map(float, y)), Fraction(0))
        got = oracle.exact_sum_diff(x, y)
        assert close(got, ex), (trial, got, float(ex))
        if ex != 0:
            assert not close(-got, ex)
    assert oracle.exact_sum_diff(np.array([1e300, -1e300, 2.0**-1074]), np.array([2.0**-1074])) == 0.0
