"""The oracle's gemm! / syrk! / gemmt! / trsm! restatements, pinned on the CPU
by mathematics rather than by the GPU library:

* integer data (hi limb only, magnitudes far below 2^53): every product and
  partial sum is exact in binary64, so any correct order yields the exact
  integer result -- pins the index maps, the transposition of syrk's P'P and
  gemmt's L R', alpha / beta, beta == 0 not reading the output (NaN there),
  the untouched upper triangle, and for trsm! all 16 side / uplo / trans /
  diag variants (power-of-two diagonals keep the divisions exact);
* multi-limb data against the exact long accumulator: the result equals the
  exact alpha*op(A)*op(B) + beta*C to within c*k*u_N of the terms' magnitude
  (gemm.jl:162-181 epilogue, exact for power-of-two alpha / beta), and every
  trsm! variant has an exact residual op(A) X - B within c*n*u_N
  (trsm.jl:18-208). A dropped limb product or a flipped sign errs by ~2^-53
  of the terms, far outside these bounds for N >= 2.
"""
import itertools

import numpy as np
import pytest

from mfutil import eps, rand_mf, case_seed


def planar(M, N):
    """dense (rows, cols) float64 -> planar (N, cols, rows), lower limbs 0."""
    out = np.zeros((N, M.shape[1], M.shape[0]))
    out[0] = M.T
    return out


def dense(P):
    return P[0].T


def ints(rng, shape, lo=-20, hi=21):
    return rng.integers(lo, hi, shape).astype(np.float64)


@pytest.mark.parametrize("N", [1, 2, 3, 4])
@pytest.mark.parametrize("m,k,n", [(7, 5, 9), (1, 13, 3), (16, 17, 15), (33, 4, 2)])
def test_gemm_exact_on_integers(oracle, N, m, k, n):
    rng = np.random.default_rng(100 * N + m + k + n)
    A, B, C = ints(rng, (m, k)), ints(rng, (k, n)), ints(rng, (m, n))
    for alpha, beta in ((1.0, 0.0), (2.0, -3.0), (-1.0, 1.0), (0.5, 0.0)):
        for fused in ((True, False) if N == 3 else (True,)):
            C0 = planar(C, N)
            if beta == 0.0:
                C0[:] = np.nan
            oracle.gemm(C0, planar(A, N), planar(B, N), alpha, beta, fused=fused)
            ref = alpha * (A @ B) + (beta * C if beta != 0.0 else 0.0)
            assert np.array_equal(dense(C0), ref), (N, alpha, beta, fused)
            assert not C0[1:].any()


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_syrk_gemmt_exact_on_integers(oracle, N):
    rng = np.random.default_rng(200 + N)
    n, k = 13, 6
    P, L, R, C = ints(rng, (k, n)), ints(rng, (n, k)), ints(rng, (n, k)), ints(rng, (n, n))
    upper = np.triu_indices(n, 1)
    for alpha, beta in ((1.0, 0.0), (-2.0, 3.0)):
        for name in ("syrk", "gemmt"):
            C0 = planar(C, N)
            C0[:, upper[1], upper[0]] = 7.25          # must survive
            if beta == 0.0:
                lower = np.tril_indices(n)
                C0[:, lower[1], lower[0]] = np.nan    # never read
            if name == "syrk":
                oracle.syrk(C0, planar(P, N), alpha, beta)
                full = alpha * (P.T @ P)
            else:
                oracle.gemmt(C0, planar(L, N), planar(R, N), alpha, beta)
                full = alpha * (L @ R.T)
            if beta != 0.0:
                full = full + beta * C
            got = dense(C0)
            il = np.tril_indices(n)
            assert np.array_equal(got[il], full[il]), (name, N, alpha, beta)
            assert np.all(got[upper] == 7.25) and np.all(C0[1:, upper[1], upper[0]] == 7.25)
            assert not C0[1:, il[1], il[0]].any()


def _aug_identity(A, beta=1.0):
    """[A | beta I] (planar), so that [A | beta I] [B; C] = A B + beta C
    exactly for a power-of-two beta."""
    N, k, m = A.shape
    I = np.zeros((N, m, m))
    I[0, np.arange(m), np.arange(m)] = beta
    return np.ascontiguousarray(np.concatenate([A, I], axis=1))


@pytest.mark.parametrize("N", [2, 3, 4])
def test_gemm_multilimb_against_exact_accumulator(oracle, N):
    rng = np.random.default_rng(300 + N)
    m, k, n = 11, 40, 7
    A, B, C = rand_mf(rng, N, m, k, spread=20), rand_mf(rng, N, k, n, spread=20), rand_mf(rng, N, m, n)
    for alpha, beta in ((1.0, 0.0), (1.0, 1.0), (0.5, 2.0), (-1.0, 1.0)):
        for fused in ((True, False) if N == 3 else (True,)):
            out = oracle.gemm(C.copy(), A, B, alpha, beta, fused=fused)
            # exact target alpha*A*B + beta*C as one product [alpha A | beta I] [B; C]
            Aa = _aug_identity(alpha * A, beta)
            Bb = np.ascontiguousarray(np.concatenate([B, C], axis=2))
            rows, cols = np.repeat(np.arange(m), n), np.tile(np.arange(n), m)
            val, scale = oracle.exact_axpb(Aa, Bb, out, sab=-1, sc=1, rows=rows, cols=cols)
            assert np.max(np.abs(val) / scale) <= 8 * (k + 1) * eps(N), (N, alpha, beta, fused)
            # the same bound rejects a result whose lowest limb is lost
            cut = out.copy()
            cut[N - 1] = 0.0
            val_c, _ = oracle.exact_axpb(Aa, Bb, cut, sab=-1, sc=1, rows=rows, cols=cols)
            assert np.max(np.abs(val_c) / scale) > 8 * (k + 1) * eps(N)


@pytest.mark.parametrize("N", [2, 3, 4])
def test_syrk_gemmt_multilimb_against_exact_accumulator(oracle, N):
    rng = np.random.default_rng(400 + N)
    n, k = 12, 30
    P = rand_mf(rng, N, k, n)
    L, R = rand_mf(rng, N, n, k), rand_mf(rng, N, n, k)
    C = rand_mf(rng, N, n, n)
    rows, cols = np.tril_indices(n)
    for name in ("syrk", "gemmt"):
        out = C.copy()
        if name == "syrk":
            oracle.syrk(out, P, 1.0, 1.0)
            left, right = np.ascontiguousarray(np.transpose(P, (0, 2, 1))), P      # P' (n x k) and P (k x n)
        else:
            oracle.gemmt(out, L, R, 1.0, 1.0)
            left, right = L, np.ascontiguousarray(np.transpose(R, (0, 2, 1)))      # L (n x k) and R' (k x n)
        Aa = _aug_identity(left)
        Bb = np.ascontiguousarray(np.concatenate([right, C], axis=2))
        val, scale = oracle.exact_axpb(Aa, Bb, out, sab=-1, sc=1, rows=rows, cols=cols)
        assert np.max(np.abs(val) / scale) <= 8 * (k + 1) * eps(N), name


SIDES, UPLOS, TRANS, DIAGS = ("left", "right"), ("lower", "upper"), ("N", "T"), ("nonunit", "unit")
VARIANTS = list(itertools.product(SIDES, UPLOS, TRANS, DIAGS))


def op_dense(A, uplo, trans, diag):
    T = np.tril(A) if uplo == "lower" else np.triu(A)
    if diag == "unit":
        np.fill_diagonal(T, 1.0)
    return T.T if trans == "T" else T


@pytest.mark.parametrize("side,uplo,trans,diag", VARIANTS)
def test_trsm_exact_on_integers(oracle, side, uplo, trans, diag):
    """op(A) X = alpha B with integer X and a triangle whose divisions are exact."""
    rng = np.random.default_rng(case_seed(side, uplo, trans, diag))
    n, r = 10, 4
    A = ints(rng, (n, n), -3, 4)
    d = 2.0 ** rng.integers(-2, 3, n)
    A[np.arange(n), np.arange(n)] = d * rng.choice([-1.0, 1.0], n)
    inactive = np.triu_indices(n, 1) if uplo == "lower" else np.tril_indices(n, -1)
    A_poison = A.copy()
    A_poison[inactive] = np.nan                                   # other triangle never read
    if diag == "unit":
        A_poison[np.arange(n), np.arange(n)] = np.nan            # nor the diagonal
    opA = op_dense(A, uplo, trans, diag)
    X = ints(rng, (n, r) if side == "left" else (r, n), -5, 6)
    Bx = opA @ X if side == "left" else X @ opA                  # exact integers
    for N in (1, 2, 4):
        for alpha in (1.0, 2.0):
            Bp = planar(Bx / alpha, N)
            oracle.trsm(Bp, planar(A_poison, N), alpha, side, uplo, trans, diag)
            assert np.array_equal(dense(Bp), X), (N, alpha)
            assert not Bp[1:].any()


@pytest.mark.parametrize("side,uplo,trans,diag", VARIANTS)
@pytest.mark.parametrize("N", [2, 4])
def test_trsm_multilimb_exact_residual(oracle, side, uplo, trans, diag, N):
    rng = np.random.default_rng(case_seed(side, uplo, trans, diag, N))
    n, r = 24, 5
    A = rand_mf(rng, N, n, n)
    if diag == "unit":
        A *= 1.0 / n                                              # keep op(A) well conditioned
    idx = np.arange(n)
    A[0, idx, idx] = np.abs(A[0, idx, idx]) + 2.0
    A[1:, idx, idx] = 0.0
    mask = np.tril(np.ones((n, n))) if uplo == "lower" else np.triu(np.ones((n, n)))
    if diag == "unit":
        np.fill_diagonal(mask, 0.0)
    At = A * mask.T[None]                                         # planar [col, row] of the used triangle
    if diag == "unit":
        At[0, idx, idx] = 1.0
    opA = np.ascontiguousarray(np.transpose(At, (0, 2, 1))) if trans == "T" else At
    B = rand_mf(rng, N, n, r) if side == "left" else rand_mf(rng, N, r, n)
    X = oracle.trsm(B.copy(), A, 1.0, side, uplo, trans, diag)
    rows, cols = np.repeat(np.arange(n), r), np.tile(np.arange(r), n)
    if side == "left":
        val, scale = oracle.exact_axpb(opA, X, B, sab=-1, sc=1, rows=rows, cols=cols)
    else:                                                         # (X op(A))' = op(A)' X'
        T = lambda M: np.ascontiguousarray(np.transpose(M, (0, 2, 1)))  # noqa: E731
        val, scale = oracle.exact_axpb(T(opA), T(X), T(B), sab=-1, sc=1, rows=rows, cols=cols)
    assert np.max(np.abs(val) / scale) <= 8 * n * eps(N)
    # the bound is tight enough to see a solve against the wrong triangle
    other = "upper" if uplo == "lower" else "lower"
    Xw = oracle.trsm(B.copy(), A + np.transpose(A, (0, 2, 1)), 1.0, side, other, trans, diag)
    if side == "left":
        val_w, _ = oracle.exact_axpb(opA, Xw, B, sab=-1, sc=1, rows=rows, cols=cols)
    else:
        val_w, _ = oracle.exact_axpb(T(opA), T(Xw), T(B), sab=-1, sc=1, rows=rows, cols=cols)
    assert np.max(np.abs(val_w) / scale) > 1e-6
