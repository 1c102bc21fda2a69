"""Ozaki-split INT8 matrix-core products (gemm_strategy = :ozaki, and the
refinement's residual updates). No MFLA counterpart and a different rounding
path from gemm!'s multiply-accumulate networks, so parity is a tolerance:
every sampled element within 2^-(52N-7) * (K max|A(i,:)| max|B(:,j)| + |beta C|)
of the exact value (the oracle's long accumulator), the digit split exact to
2^-(7S-8) relative to the row maximum (product with the identity), and the
refined solve's true backward error at the Float64xN floor."""
import numpy as np
import pytest
import torch

from mfutil import rand_mf

pytestmark = pytest.mark.gpu


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def rowmax(A):          # planar (N, cols, rows) -> max |hi| per row
    return np.abs(A[0]).max(axis=0)


def colmax(B):
    return np.abs(B[0]).max(axis=1)


@pytest.mark.parametrize("N", [2, 3, 4])
@pytest.mark.parametrize("m,k,n", [(100, 130, 37), (256, 512, 128), (64, 8192, 48)])
def test_ozaki_gemm_accuracy(oracle, N, m, k, n):
    from mfla import api as M

    rng = np.random.default_rng(m + k + n + N)
    A = rand_mf(rng, N, m, k, spread=6)
    B = rand_mf(rng, N, k, n, spread=6)
    C0 = rand_mf(rng, N, m, n)
    cfg = M.KernelConfig(gemm_strategy="ozaki")
    rows = rng.integers(0, m, 300)
    cols = rng.integers(0, n, 300)
    pad = lambda X: np.concatenate([X, np.zeros_like(X)], axis=0)          # 2N limbs, same value
    A2, B2 = pad(A), pad(B)
    for sab, sc in ((1, 0), (-1, 1), (1, -1)):     # the accumulator takes unit signs
        C = cu(C0)
        M.gemm_(C, cu(A), cu(B), float(sab), float(sc), config=cfg)
        # exact  sab*A*B - (C_gpu - sc*C0):  the "C" operand carries C_gpu's
        # limbs followed by -sc*C0's limbs (the accumulator sums all limbs)
        Cmix = np.concatenate([C.cpu().numpy(), -sc * C0], axis=0)
        err, _ = oracle.exact_axpb(A2, B2, Cmix, sab=sab, sc=-1, rows=rows, cols=cols)
        tol = 2.0 ** (-52 * N + 7) * (k * rowmax(A)[rows] * colmax(B)[cols] * abs(sab) + abs(sc) * np.abs(C0[0][cols, rows]))
        assert np.all(np.abs(err) <= tol), (np.max(np.abs(err) / tol), N, m, k, n, sab, sc)


@pytest.mark.parametrize("N", [2, 3, 4])
def test_ozaki_split_exact_through_identity(oracle, N):
    from mfla import api as M

    rng = np.random.default_rng(40 + N)
    m, n = 96, 80
    A = rand_mf(rng, N, m, n, spread=20)
    I = np.zeros((N, n, n))
    I[0, np.arange(n), np.arange(n)] = 1.0
    C = torch.zeros((N, n, m), dtype=torch.float64, device="cuda")
    M.gemm_(C, cu(A), cu(I), config=M.KernelConfig(gemm_strategy="ozaki"))
    rows = np.repeat(np.arange(m), n)
    cols = np.tile(np.arange(n), m)
    val, _ = oracle.exact_axpb(A, I, C.cpu().numpy(), sab=1, sc=-1, rows=rows, cols=cols)   # A - C exactly
    S = {2: 19, 3: 27, 4: 34}[N]
    bound = np.maximum(rowmax(A)[rows] * 2.0 ** (-7 * S + 8), 2.0 ** (-52 * N + 2) * rowmax(A)[rows])
    assert np.all(np.abs(val) <= bound)


def test_ozaki_gemm_zero_rows_and_limits(oracle):
    from mfla import api as M

    N = 3
    rng = np.random.default_rng(9)
    A = rand_mf(rng, N, 50, 70)
    A[:, :, 7] = 0.0                                  # row 7 of A is zero
    B = rand_mf(rng, N, 70, 30)
    C = cu(np.full((N, 30, 50), np.nan))
    M.gemm_(C, cu(A), cu(B), config=M.KernelConfig(gemm_strategy="ozaki"))       # beta = 0: C not read
    Cg = C.cpu().numpy()
    assert np.all(np.isfinite(Cg)) and np.all(Cg[:, :, 7] == 0.0)
    with pytest.raises(M.ArgumentError):                 # reduction length 8200 > 8192
        M.gemm_(torch.zeros((N, 4, 4), dtype=torch.float64, device="cuda"),
                torch.zeros((N, 8200, 4), dtype=torch.float64, device="cuda"),
                torch.zeros((N, 4, 8200), dtype=torch.float64, device="cuda"), config=M.KernelConfig(gemm_strategy="ozaki"))


@pytest.mark.parametrize("kind,uplo,n", [("cholesky", "general", 256), ("cholesky", "lower", 200), ("lu", "general", 192)])
@pytest.mark.parametrize("N", [2, 3, 4])
def test_refinement_with_ozaki_residuals(oracle, kind, uplo, n, N):
    from mfla import refine as R
    from test_gpu_refine import exact_backward_error

    nrhs = 4
    A = R.spd_kappa(n, 1e8, N=N) if kind == "cholesky" else R.general_kappa(n, 1e8, N=N)
    if uplo == "lower":
        iu = torch.triu_indices(n, n, 1, device="cuda")
        A = A.clone()
        A[:, iu[1], iu[0]] = float("nan")
        A = A.contiguous()
    rng = np.random.default_rng(15 + N)
    Bh = np.zeros((N, nrhs, n))
    Bh[0] = rng.standard_normal((nrhs, n))
    for l in range(1, N):
        Bh[l] = Bh[l - 1] * 2.0**-53 * rng.uniform(-1, 1, (nrhs, n))
    X, hist = R.refine_solve(A, cu(Bh), kind, uplo=uplo, residual="ozaki")
    assert 2 <= len(hist) <= 12, hist
    Ah = A.cpu().numpy()
    if uplo == "lower":
        il = np.triu_indices(n, 1)
        Ah[:, il[1], il[0]] = Ah[:, il[0], il[1]]
    be = exact_backward_error(oracle, Ah, X.cpu().numpy(), Bh)
    assert be.max() <= 2.0 ** (-52 * N + 8), (be.max(), hist)


@pytest.mark.parametrize("kind,n,nrhs", [("cholesky", 1000, 64), ("lu", 777, 33)])
def test_refinement_with_ozaki_residuals_many_rhs(oracle, kind, n, nrhs):
    """Ragged row count (not a multiple of the 16-row output blocks) and up to
    64 right-hand sides through the narrow digit product."""
    from mfla import refine as R
    from test_gpu_refine import exact_backward_error

    N = 4
    A = R.spd_kappa(n, 1e8, N=N) if kind == "cholesky" else R.general_kappa(n, 1e8, N=N)
    rng = np.random.default_rng(29)
    Bh = np.zeros((N, nrhs, n))
    Bh[0] = rng.standard_normal((nrhs, n))
    for l in range(1, N):
        Bh[l] = Bh[l - 1] * 2.0**-53 * rng.uniform(-1, 1, (nrhs, n))
    X, hist = R.refine_solve(A, cu(Bh), kind, residual="ozaki")
    assert 2 <= len(hist) <= 12, hist
    be = exact_backward_error(oracle, A.cpu().numpy(), X.cpu().numpy(), Bh)
    assert be.max() <= 2.0 ** (-52 * N + 8), (be.max(), hist)
