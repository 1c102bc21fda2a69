"""FP64-factor + Float64x4 refinement (a build extension: MFLA has no FP64 ->
x4 refinement loop). Pinned by the exact long accumulator: the refined
solution's true normwise backward error (residual evaluated exactly, not in
Float64x4) must be <= 1e-60, like the plain Float64x4 factor + solve."""
import numpy as np
import pytest
import torch

from mfutil import HELD_OFF_GPU

pytestmark = pytest.mark.gpu


def exact_backward_error(oracle, A, X, B):
    N, n, m = A.shape
    nrhs = X.shape[1]
    rows = np.repeat(np.arange(m), nrhs)
    cols = np.tile(np.arange(nrhs), m)
    val, _ = oracle.exact_axpb(A, X, B, sab=-1, sc=1, rows=rows, cols=cols)
    rn = np.abs(val).reshape(m, nrhs).max(axis=0)
    an = np.abs(A.sum(0)).sum(axis=0).max()            # row sums of |A| (A.sum(0) is [col, row])
    xn = np.abs(X.sum(0)).max(axis=1)
    bn = np.abs(B.sum(0)).max(axis=1)
    return rn / (an * xn + bn)


@pytest.mark.parametrize("kind", ["cholesky", "lu"])
def test_refined_solve_reaches_x4_backward_error(oracle, kind):
    from mfla import refine as R

    n, nrhs = 256, 4
    A = R.spd_kappa(n, 1e8, N=4) if kind == "cholesky" else R.general_kappa(n, 1e8, N=4)
    rng = np.random.default_rng(5)
    Bh = np.zeros((4, nrhs, n))
    Bh[0] = rng.standard_normal((nrhs, n))
    B = torch.from_numpy(Bh).cuda()
    X, hist = R.refine_solve(A, B, kind)
    assert len(hist) >= 3 and hist[-1] <= 1e-60, hist
    # the convergence is geometric at ~c kappa * u_64 per step (dense Haar
    # factors: kappa * u_64 ~ 1e-8 times a modest growth constant)
    assert hist[1] < 1e-4 * hist[0]
    Ah, Xh = A.cpu().numpy(), X.cpu().numpy()
    be = exact_backward_error(oracle, Ah, Xh, Bh)
    assert be.max() <= 1e-60, be
    Xp = R.plain_solve(A, B, kind).cpu().numpy()
    bp = exact_backward_error(oracle, Ah, Xp, Bh)
    assert bp.max() <= 1e-60, bp


def _np(t):
    return t.cpu().numpy()


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (64, 64, 64), (100, 37, 130), (257, 70, 65)])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_dgemm_matches_float64_reference(m, n, k, ta, tb):
    from mfla import api as M

    rng = np.random.default_rng(m * 7 + n * 3 + k)
    A = rng.standard_normal((k, m) if ta == "N" else (m, k))     # planar [col, row]
    B = rng.standard_normal((n, k) if tb == "N" else (k, n))
    C0 = rng.standard_normal((n, m))
    opA = A.T if ta == "N" else A                                 # rows x cols views
    opB = B.T if tb == "N" else B
    for alpha, beta in ((1.0, 0.0), (-0.5, 2.0)):
        C = torch.from_numpy(C0.copy()[None]).cuda()
        M.dgemm_(C, torch.from_numpy(A[None]).cuda(), torch.from_numpy(B[None]).cuda(), alpha, beta, trans_a=ta, trans_b=tb)
        ref = alpha * opA @ opB + beta * C0.T
        err = np.abs(_np(C)[0].T - ref).max() / max(1.0, np.abs(ref).max())
        assert err <= 4 * k * 2.0**-52, (m, n, k, ta, tb, err)


@pytest.mark.parametrize("tri_a,tri_b", [("lower", "general"), ("upper", "general"), ("general", "lower"), ("general", "upper")])
def test_dgemm_triangular_operands_ignore_the_other_triangle(tri_a, tri_b):
    from mfla import api as M

    rng = np.random.default_rng(11)
    n, r = 150, 40
    A = rng.standard_normal((n, n))
    B = rng.standard_normal((r, n)) if tri_b == "general" else rng.standard_normal((n, n))
    Bm = B.T
    Am = A.T
    if tri_a != "general":
        keep = np.tril(np.ones((n, n))) if tri_a == "lower" else np.triu(np.ones((n, n)))
        Aref = Am * keep
        Am_poison = np.where(keep > 0, Am, np.nan)
        A = Am_poison.T.copy()
        Bref = Bm
        nout = r
    else:
        keep = np.tril(np.ones((n, n))) if tri_b == "lower" else np.triu(np.ones((n, n)))
        Aref = Am
        Bref = Bm * keep
        B = np.where(keep > 0, Bm, np.nan).T.copy()
        nout = n
    C = torch.zeros((1, nout, n), dtype=torch.float64, device="cuda")
    M.dgemm_(C, torch.from_numpy(A[None]).cuda(), torch.from_numpy(B[None]).cuda(), tri_a=tri_a, tri_b=tri_b)
    ref = Aref @ Bref
    assert np.isfinite(_np(C)).all()
    assert np.abs(_np(C)[0].T - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("n", [1, 5, 63, 64, 65, 129, 130, 200, 300, 1000])
def test_dpotrf_inverse(n):
    from mfla import api as M

    rng = np.random.default_rng(n)
    G = rng.standard_normal((n, n))
    S = G @ G.T + n * np.eye(n)
    Sp = np.where(np.tril(np.ones((n, n))) > 0, S, np.nan)     # the upper triangle is never read
    A = torch.from_numpy(Sp.T.copy()[None]).cuda()
    Inv = torch.full((1, n, n), np.nan, dtype=torch.float64, device="cuda")
    assert M.dpotrf_inverse_(A, Inv) == 0
    L = np.tril(_np(A)[0].T)
    Li = np.tril(_np(Inv)[0].T)
    Lref = np.linalg.cholesky(S)
    assert np.abs(L - Lref).max() <= 1e-13 * np.abs(Lref).max() * max(1, n ** 0.5)
    assert np.abs(Li @ L - np.eye(n)).max() <= 1e-13 * n


def test_dpotrf_inverse_reports_failing_pivot():
    from mfla import api as M

    n = 100
    S = np.eye(n) * 4.0
    S[70, 70] = -1.0
    A = torch.from_numpy(S.T.copy()[None]).cuda()
    Inv = torch.zeros_like(A)
    assert M.dpotrf_inverse_(A, Inv) == 71
    S[3, 3] = np.inf
    A = torch.from_numpy(S.T.copy()[None]).cuda()
    assert M.dpotrf_inverse_(A, Inv) == -1


def _lu_check(A0, F, ipiv, pivots_like_lapack):
    import scipy.linalg as sl

    n = A0.shape[0]
    PA = A0.copy()
    for k in range(n):
        p = ipiv[k] - 1
        assert k <= p < n
        PA[[k, p]] = PA[[p, k]]
    Lf = np.tril(F, -1) + np.eye(n)
    U = np.triu(F)
    assert np.abs(np.tril(F, -1)).max(initial=0.0) <= 1.0
    assert np.abs(PA - Lf @ U).max() <= 1e-15 * n * np.abs(A0).max() * 4
    if pivots_like_lapack:
        _, piv = sl.lu_factor(A0)
        assert np.array_equal(ipiv - 1, piv)


@pytest.mark.parametrize("n", [1, 5, 8, 17, 64, 65, 129, 130, 200, 300, 1100, 2100, 4100])
def test_dgetrf_blocked_lu(n):
    from mfla import api as M

    rng = np.random.default_rng(100 + n)
    A0 = rng.standard_normal((n, n))
    A = torch.from_numpy(A0.T.copy()[None]).cuda()
    ipiv = torch.zeros(n, dtype=torch.int64, device="cuda")
    assert M.dgetrf_(A, ipiv) == 0
    _lu_check(A0, _np(A)[0].T, ipiv.cpu().numpy(), pivots_like_lapack=n <= 300)


def test_dgetrf_zero_pivot_and_nonfinite():
    from mfla import api as M

    n = 100
    rng = np.random.default_rng(5)
    A0 = rng.standard_normal((n, n))
    A0[:, 37] = 0.0                                   # inside the third 16-column leaf
    A = torch.from_numpy(A0.T.copy()[None]).cuda()
    ipiv = torch.zeros(n, dtype=torch.int64, device="cuda")
    assert M.dgetrf_(A, ipiv) == 38
    assert ipiv.cpu().numpy()[37] == 38
    A0[3, 3] = np.nan
    A = torch.from_numpy(A0.T.copy()[None]).cuda()
    assert M.dgetrf_(A, ipiv) == -1


@pytest.mark.parametrize("kind,uplo,n,nrhs", [("cholesky", "general", 256, 4), ("cholesky", "lower", 200, 3),
                                              ("lu", "general", 192, 5)])
@pytest.mark.parametrize("N", [2, 3, 4])
def test_native_refinement_reaches_working_precision(oracle, kind, uplo, n, nrhs, N):
    from mfla import refine as R

    A = R.spd_kappa(n, 1e8, N=N) if kind == "cholesky" else R.general_kappa(n, 1e8, N=N)
    if uplo == "lower":                     # poison the triangle that must not be read
        iu = torch.triu_indices(n, n, 1, device="cuda")
        A = A.clone()
        A[:, iu[1], iu[0]] = float("nan")  # planar [col, row]: col > row is the upper triangle
        A = A.contiguous()
    rng = np.random.default_rng(5 + N)
    Bh = np.zeros((N, nrhs, n))
    Bh[0] = rng.standard_normal((nrhs, n))
    for l in range(1, N):
        Bh[l] = Bh[l - 1] * 2.0**-53 * rng.uniform(-1, 1, (nrhs, n))
    B = torch.from_numpy(Bh).cuda()
    X, hist = R.refine_solve(A, B, kind, uplo=uplo, residual="limbs")
    assert 2 <= len(hist) <= 12 and hist[-1] <= 64 * 2.0 ** (-52 * N), hist
    Ah = A.cpu().numpy()
    if uplo == "lower":
        il = np.triu_indices(n, 1)
        Ah[:, il[1], il[0]] = Ah[:, il[0], il[1]]      # symmetric from the lower triangle
    be = exact_backward_error(oracle, Ah, X.cpu().numpy(), Bh)
    assert be.max() <= 2.0 ** (-52 * N + 8), (be.max(), hist)


def test_native_refinement_errors():
    from mfla import api as M
    from mfla import refine as R

    n = 64
    A = torch.zeros((4, n, n), dtype=torch.float64, device="cuda")
    A[0] = -torch.eye(n, dtype=torch.float64, device="cuda")
    B = torch.ones((4, 2, n), dtype=torch.float64, device="cuda")
    with pytest.raises(M.PosDefException):
        R.refine_solve(A, B, "cholesky")
    with pytest.raises(M.DimensionMismatch):
        R.refine_solve(A, B[:, :, :10].contiguous(), "cholesky")
    with pytest.raises(M.ArgumentError):
        R.refine_solve(A[:1].contiguous(), B[:1].contiguous(), "cholesky")


@HELD_OFF_GPU
def test_refinement_without_convergence_keeps_the_best_iterate():
    """kappa = 1e17 > 1 / u_64: the binary64 LU factor still exists but the
    corrections stop contracting. check=True raises ConvergenceError, and with
    check=False the returned X is the best iterate (a correction that made
    the estimate worse is taken back)."""
    from mfla import api as M
    from mfla import refine as R

    n, nrhs = 256, 2
    A = R.general_kappa(n, 1e17, N=4)
    rng = np.random.default_rng(17)
    Bh = np.zeros((4, nrhs, n))
    Bh[0] = rng.standard_normal((nrhs, n))
    B = torch.from_numpy(Bh).cuda()
    with pytest.raises(M.ConvergenceError) as ei:
        R.refine_solve(A, B, "lu", residual="limbs")
    assert ei.value.X is not None and min(ei.value.history) > 2.0 ** -216
    X, hist = R.refine_solve(A, B, "lu", residual="limbs", check=False)
    Rr = torch.zeros_like(B)
    M.residual_(Rr, A, X, B)
    be = float(np.max(np.abs(M.normwise_backward_error(A, X, B, Rr)[0])))
    assert be <= 4 * min(hist), (be, hist)


@pytest.mark.parametrize("bad", [70, 200, 299])
def test_dgetrf_zero_column_in_later_blocks(bad):
    """A zero column in a later 64-column block (and in the last, partial one):
    info names it, its pivot is the identity, and the factors of the columns
    before it still reproduce P A."""
    from mfla import api as M

    n = 300
    rng = np.random.default_rng(bad)
    A0 = rng.standard_normal((n, n))
    A0[:, bad] = 0.0
    A = torch.from_numpy(A0.T.copy()[None]).cuda()
    ipiv = torch.zeros(n, dtype=torch.int64, device="cuda")
    assert M.dgetrf_(A, ipiv) == bad + 1
    piv = ipiv.cpu().numpy()
    assert piv[bad] == bad + 1
    assert all(k + 1 <= p <= n for k, p in enumerate(piv))
    F = _np(A)[0].T
    PA = A0.copy()
    for k in range(n):
        p = piv[k] - 1
        PA[[k, p]] = PA[[p, k]]
    Lf = np.tril(F, -1) + np.eye(n)
    U = np.triu(F)
    c = bad                                       # columns before the zero one are fully factored
    assert np.abs(PA[:, :c] - (Lf @ U)[:, :c]).max() <= 1e-13 * n


@pytest.mark.parametrize("bad", [3, 40, 104, 250])
def test_dpotrf_inverse_failing_pivot_in_later_panels(bad):
    """Failing pivots in each 16-column panel of a 64 x 64 diagonal block and
    in a later block report info = column + 1."""
    from mfla import api as M

    n = 300
    rng = np.random.default_rng(bad)
    G = rng.standard_normal((n, n))
    S = G @ G.T / n + np.eye(n)
    S[bad, bad] = -5.0
    A = torch.from_numpy(S.T.copy()[None]).cuda()
    Inv = torch.zeros_like(A)
    info = M.dpotrf_inverse_(A, Inv)
    # every leading block before `bad` is SPD and the Schur complement pivot
    # at `bad` is -5 minus a nonnegative term
    np.linalg.cholesky(S[:bad, :bad])
    assert info == bad + 1
