"""Pins the oracle's multi-double networks (CPU only).

* x3 standard (acc + x*y) and fused mulacc_x3 networks against the eight
  fixed limb vectors of test/mulacc_x3_proof_vectors.jl:1-58 (bitwise).
* TwoSum / FastTwoSum / TwoProd error-freeness (s + e == a + b exactly).
* x2/x3/x4 add, mul, div, sqrt against 512-bit exact values (committed
  fixture tests/golden/arith_exact.json, written by gen_arith_goldens.py, and
  live mpmath on fresh seeds): relative error <= 2 eps(Float64xN).
* algebraic invariants: commutativity, normalised outputs, exact cases.
"""
import json
import os

import mpmath
import numpy as np
import pytest

from mfutil import eps, rand_mf

HERE = os.path.dirname(os.path.abspath(__file__))
mpmath.mp.prec = 600


def _f(bits):
    return np.array(bits, dtype=np.uint64).view(np.float64)


def _proof_vectors():
    with open(os.path.join(HERE, "golden", "mulacc_x3_proof_vectors.json")) as f:
        pv = json.load(f)
    acc = np.stack([_f(v["acc"]) for v in pv], axis=1)
    x = np.stack([_f(v["x"]) for v in pv], axis=1)
    y = np.stack([_f(v["y"]) for v in pv], axis=1)
    out = np.stack([_f(v["output"]) for v in pv], axis=1)
    return acc, x, y, out


def test_proof_vectors_standard_and_fused(oracle):
    acc, x, y, out = _proof_vectors()
    std = oracle.arith("mac", x, y, acc)
    fused = oracle.arith("mulacc_x3", x, y, acc)
    assert np.array_equal(std.view(np.uint64), out.view(np.uint64))
    assert np.array_equal(fused.view(np.uint64), out.view(np.uint64))


def test_proof_vectors_catch_a_dropped_term(oracle):
    # A plausible restatement mistake -- dropping the x2*y0 partial product
    # from the x3 prefix -- changes the committed outputs: the vectors pin it.
    acc, x, y, out = _proof_vectors()
    x_trunc = x.copy()
    x_trunc[2] = 0.0
    bad = oracle.arith("mac", x_trunc, y, acc)
    assert not np.array_equal(bad.view(np.uint64), out.view(np.uint64))


def _mp(x, i):
    return mpmath.fsum(mpmath.mpf(float(v)) for v in x[:, i])


def test_eft_error_free(oracle):
    rng = np.random.default_rng(7)
    a = rng.standard_normal(2000) * 2.0 ** rng.integers(-60, 60, 2000)
    b = rng.standard_normal(2000) * 2.0 ** rng.integers(-60, 60, 2000)
    b[:100] = -a[:100] * (1 + 2.0**-40)
    s, e = oracle.two_sum(a, b)
    p, pe = oracle.two_prod(a, b)
    big = np.abs(a) >= np.abs(b)
    fs, fe = oracle.two_sum(np.where(big, a, b), np.where(big, b, a), fast=True)
    for i in range(0, 2000, 7):
        A, B = mpmath.mpf(float(a[i])), mpmath.mpf(float(b[i]))
        assert mpmath.mpf(float(s[i])) + mpmath.mpf(float(e[i])) == A + B
        assert s[i] == a[i] + b[i]
        assert mpmath.mpf(float(fs[i])) + mpmath.mpf(float(fe[i])) == A + B
        assert fs[i] == s[i] and fe[i] == e[i]          # ordered FastTwoSum == TwoSum
        assert mpmath.mpf(float(p[i])) + mpmath.mpf(float(pe[i])) == A * B


def test_golden_exact_values(oracle):
    with open(os.path.join(HERE, "golden", "arith_exact.json")) as f:
        g = json.load(f)
    mpmath.mp.prec = 512
    worst = {}
    for c in g["cases"]:
        N, op = c["N"], c["op"]
        x = _f(c["x"]).reshape(N, 1)
        y = _f(c["y"]).reshape(N, 1)
        r = oracle.arith(op, x, None if op == "sqrt" else y)
        got = _mp(r, 0)
        ex = mpmath.mpf(c["exact"])
        scale = mpmath.mpf(c["scale"])
        rel = float(abs(got - ex) / scale) if scale != 0 else 0.0
        worst[(N, op)] = max(worst.get((N, op), 0.0), rel)
        assert rel <= 2 * eps(N), (N, op, rel)
    assert len(worst) == 12


@pytest.mark.parametrize("N", [2, 3, 4])
def test_live_accuracy_and_invariants(oracle, N):
    rng = np.random.default_rng(100 + N)
    cnt = 400
    a = rand_mf(rng, N, cnt, spread=40)
    b = rand_mf(rng, N, cnt, spread=40)
    b[:, :40] = -a[:, :40]
    b[N - 1, :40] += a[N - 1, :40] * 2.0**-9
    s = oracle.arith("add", a, b)
    p = oracle.arith("mul", a, b)
    # add is commutative bitwise; the x2 / x4 products fuse their cross terms
    # into FMAs in operand order, so mul(b, a) may differ from mul(a, b) in
    # the last limb (the x3 product, MFLA's own network, is commutative) --
    # both orders are within the accuracy bound below, and every kernel
    # multiplies in MFLA's operand order
    assert np.array_equal(s.view(np.uint64), oracle.arith("add", b, a).view(np.uint64))
    pr = oracle.arith("mul", b, a)
    if N == 3:
        assert np.array_equal(p.view(np.uint64), pr.view(np.uint64))
    for i in range(0, cnt, 5):
        assert float(abs(_mp(pr, i) - _mp(a, i) * _mp(b, i)) / abs(_mp(a, i) * _mp(b, i))) <= 2 * eps(N)
    for i in range(0, cnt, 5):
        A, B = _mp(a, i), _mp(b, i)
        assert float(abs(_mp(s, i) - (A + B)) / (abs(A) + abs(B))) <= 2 * eps(N)
        assert float(abs(_mp(p, i) - A * B) / abs(A * B)) <= 2 * eps(N)
        # normalised: each limb absorbs the next one (fl(x_k + x_{k+1}) == x_k)
        for r in (s, p):
            for l in range(N - 1):
                if r[l, i] != 0:
                    assert r[l, i] + r[l + 1, i] == r[l, i]


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_exact_cases(oracle, N):
    # products and quotients that are exactly representable come out exact
    x = np.zeros((N, 4))
    y = np.zeros((N, 4))
    x[0] = [3.0, -7.5, 2.0**-300, 1.0]
    y[0] = [0.25, 2.0**40, 2.0**300, 1.0]
    p = oracle.arith("mul", x, y)
    assert np.array_equal(p[0], x[0] * y[0]) and np.all(p[1:] == 0)
    q = oracle.arith("div", p, y)
    assert np.array_equal(q[0], x[0]) and np.all(q[1:] == 0)
    sq = oracle.arith("sqrt", np.abs(x) ** 2 * 0 + np.array([[9.0, 2.0**-600, 16.0, 1.0]] + [[0.0] * 4] * (N - 1)))
    assert np.array_equal(sq[0], [3.0, 2.0**-300, 4.0, 1.0])
    if N == 3:
        acc = rand_mf(np.random.default_rng(3), 3, 4)
        assert np.array_equal(oracle.arith("mac", x, y, acc).view(np.uint64), oracle.arith("mulacc_x3", x, y, acc).view(np.uint64))
