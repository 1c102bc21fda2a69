"""The oracle's trmm!, mfdot and syrk_packed! restatements, pinned on the CPU:

* exactness on integer data (every product and partial sum is exact in
  binary64, so any correct order gives the exact integer matrix product):
  pins the triangle, transposition, side, unit diagonal and alpha handling;
* multi-limb data against the exact long accumulator (within c*n*u_N):
  pins that the right terms enter with the right signs;
* packed SYRK equals the dense syrk! restatement entry by entry (bitwise)
  and the packed index map (r, c) -> c(2n - c + 1)/2 + (r - c).
"""
import itertools

import numpy as np
import pytest

from mfutil import assert_bitwise, eps, rand_mf, case_seed

SIDES, UPLOS, TRANS, DIAGS = ("left", "right"), ("lower", "upper"), ("N", "T"), ("nonunit", "unit")


def op_dense(A, uplo, trans, diag):
    """op(A) of the dense (rows x cols) matrix restricted to its triangle."""
    T = np.tril(A) if uplo == "lower" else np.triu(A)
    if diag == "unit":
        np.fill_diagonal(T, 1)
    return T.T if trans == "T" else T


@pytest.mark.parametrize("side,uplo,trans,diag", list(itertools.product(SIDES, UPLOS, TRANS, DIAGS)))
def test_trmm_exact_on_integers(oracle, side, uplo, trans, diag):
    rng = np.random.default_rng(case_seed(side, uplo, trans, diag))
    N, n, r = 3, 9, 5
    A = rng.integers(-9, 10, (n, n)).astype(float)
    A_poison = A.copy()
    inactive = np.triu_indices(n, 1) if uplo == "lower" else np.tril_indices(n, -1)
    A_poison[inactive] = np.nan                                   # never read
    if diag == "unit":
        A_poison[np.arange(n), np.arange(n)] = np.nan
    B = rng.integers(-9, 10, (n, r) if side == "left" else (r, n)).astype(float)
    Ap = np.zeros((N, n, n)); Ap[0] = A_poison.T
    Bp = np.zeros((N,) + B.T.shape); Bp[0] = B.T
    oracle.trmm(Bp, Ap, 2.0, side, uplo, trans, diag)
    opA = op_dense(A, uplo, trans, diag)
    ref = 2 * (opA @ B if side == "left" else B @ opA)
    assert np.array_equal(Bp[0].T, ref) and not Bp[1:].any()


@pytest.mark.parametrize("side,uplo,trans", list(itertools.product(SIDES, UPLOS, TRANS)))
@pytest.mark.parametrize("N", [2, 4])
def test_trmm_multilimb_against_exact_accumulator(oracle, side, uplo, trans, N):
    rng = np.random.default_rng(7 + N)
    n, r = 17, 6
    A = rand_mf(rng, N, n, n)
    B = rand_mf(rng, N, n, r) if side == "left" else rand_mf(rng, N, r, n)
    X = B.copy()
    oracle.trmm(X, A, 1.0, side, uplo, trans, "nonunit")
    mask = np.tril(np.ones((n, n))) if uplo == "lower" else np.triu(np.ones((n, n)))
    At = A * mask.T[None]                                         # planar [col, row]
    opA = np.ascontiguousarray(np.transpose(At, (0, 2, 1))) if trans == "T" else At
    if side == "left":
        rows, cols = np.repeat(np.arange(n), r), np.tile(np.arange(r), n)
        val, scale = oracle.exact_axpb(opA, B, X, sab=-1, sc=1, rows=rows, cols=cols)
    else:                                                         # (B op(A))^T = op(A)^T B^T
        opAT = np.ascontiguousarray(np.transpose(opA, (0, 2, 1)))
        BT = np.ascontiguousarray(np.transpose(B, (0, 2, 1)))
        XT = np.ascontiguousarray(np.transpose(X, (0, 2, 1)))
        rows, cols = np.repeat(np.arange(n), r), np.tile(np.arange(r), n)
        val, scale = oracle.exact_axpb(opAT, BT, XT, sab=-1, sc=1, rows=rows, cols=cols)
    assert np.all(np.abs(val) <= 8 * n * eps(N) * scale)


def test_dot_exact_and_accurate(oracle):
    rng = np.random.default_rng(3)
    x = np.zeros((3, 1000)); y = np.zeros((3, 1000))
    x[0] = rng.integers(-1000, 1000, 1000); y[0] = rng.integers(-1000, 1000, 1000)
    d = oracle.dot(x, y)
    assert d[0] == float(np.dot(x[0].astype(np.int64), y[0].astype(np.int64))) and not d[1:].any()
    for N in (2, 3, 4):
        X = rand_mf(rng, N, 40, 7)
        Y = rand_mf(rng, N, 40, 7)
        d = oracle.dot(X, Y)
        D = np.zeros((N, 1, 1)); D[:, 0, 0] = d
        val, scale = oracle.exact_axpb(np.ascontiguousarray(X.reshape(N, -1, 1)), np.ascontiguousarray(Y.reshape(N, 1, -1)),
                                       D, sab=-1, sc=1, rows=np.array([0]), cols=np.array([0]))
        assert abs(val[0]) <= 8 * 280 * eps(N) * scale[0]


@pytest.mark.parametrize("N", [1, 3, 4])
def test_syrk_packed_equals_dense_syrk(oracle, N):
    rng = np.random.default_rng(20 + N)
    n, k = 11, 9
    P = rand_mf(rng, N, k, n)
    for rf, rl, beta in ((1, k, 0.0), (3, 7, 0.5)):
        C = rand_mf(rng, N, n, n)
        packed = np.zeros((N, n * (n + 1) // 2))
        for c in range(n):
            for r in range(c, n):
                packed[:, c * (2 * n - c + 1) // 2 + (r - c)] = C[:, c, r]
        oracle.syrk_packed(packed, P, 1.5, beta, rf, rl)
        Pr = np.ascontiguousarray(P[:, :, rf - 1: rl])
        oracle.syrk(C, Pr, 1.5, beta)
        for c in range(n):
            for r in range(c, n):
                assert_bitwise(packed[:, c * (2 * n - c + 1) // 2 + (r - c)], C[:, c, r], f"({r},{c})")
