"""Reverse-mode rules for solve, Cholesky and logdet (MFLA has none; these are
the pullbacks of the rrules a Julia caller registers). Pinned two ways:

* against a 512-bit reference (mpmath at prec=512 on the exact values of the
  Float64xN inputs): logdet, Bbar = S^{-1} Xbar, Sbar = -sym(Bbar X^T),
  ybar S^{-1} and sym(L^{-T} Phi(L^T Lbar) L^{-1});
* against central finite differences computed in Float64xN on the device,
  h = 2^-80 for Float64x4 (truncation h^2 = 2^-160, rounding u_4/h = 2^-128)
  and h = 2^-52 for Float64x3 (both terms ~2^-104): <Xbar, dX/dh> =
  <Sbar, E> + <Bbar, E_B> for random symmetric E, checked exactly (the
  oracle's long accumulator forms the difference of the two inner products
  before rounding), so the agreement is verified to ~1e-30 (x4) / ~1e-26 (x3)
  relative, far below binary64.
"""
import mpmath
import numpy as np
import pytest
import torch

from mfutil import rand_mf

pytestmark = pytest.mark.gpu

H_EXP = {3: 52, 4: 80}      # h = 2^-H_EXP[N]
FD_TOL = {3: 1e-26, 4: 1e-30}


def spd(oracle, rng, N, n, shift=None):
    R = rand_mf(rng, N, n, n)
    S = np.zeros((N, n, n))
    oracle.gemm(S, R, np.ascontiguousarray(np.transpose(R, (0, 2, 1))), fused=False)
    S[0, np.arange(n), np.arange(n)] += float(n if shift is None else shift)
    # exactly symmetric: mirror the lower triangle
    iu = np.triu_indices(n, 1)
    S[:, iu[1], iu[0]] = S[:, iu[0], iu[1]]
    return S


def to_mp(x):
    """Planar (N, cols, rows) -> mpmath matrix of the exact limb sums."""
    N, cols, rows = x.shape
    M = mpmath.matrix(rows, cols)
    for j in range(cols):
        for i in range(rows):
            M[i, j] = mpmath.fsum(mpmath.mpf(float(x[l, j, i])) for l in range(N))
    return M


def rel_err(ours, ref):
    d = to_mp(ours) - ref
    num = max(abs(d[i, j]) for i in range(d.rows) for j in range(d.cols))
    den = max(abs(ref[i, j]) for i in range(ref.rows) for j in range(ref.cols))
    return float(num / den)


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def lower_of(F):
    N, n, _ = F.shape
    out = F.copy()
    iu = np.triu_indices(n, 1)
    out[:, iu[1], iu[0]] = 0.0          # planar [col, row]: col > row is the upper triangle
    return out


@pytest.mark.parametrize("N", [2, 3, 4])
def test_logdet_and_adjoints_match_512_bit_reference(oracle, N):
    from mfla import api as M
    from mfla import autodiff as AD

    mpmath.mp.prec = 512
    rng = np.random.default_rng(70 + N)
    n, nrhs = 12, 3
    S = spd(oracle, rng, N, n)
    B = rand_mf(rng, N, n, nrhs)
    Xbar = rand_mf(rng, N, n, nrhs)
    F = M.cholesky_(cu(S))
    X = M.solve(F, cu(B))
    Sm, Bm, Xbm = to_mp(S), to_mp(B), to_mp(Xbar)
    Sinv = Sm ** -1
    tol = 2.0 ** (-52 * N + 12)

    ld = AD.logdet(F)
    ref = mpmath.log(mpmath.det(Sm))
    assert abs(mpmath.fsum(mpmath.mpf(float(v)) for v in ld) - ref) <= tol * abs(ref)

    Bbar, Sbar = AD.solve_adjoint(F, X, cu(Xbar))
    Xm = Sinv * Bm
    Bref = Sinv * Xbm
    G = Bref * Xm.T
    Sref = -(G + G.T) / 2
    assert rel_err(Bbar.cpu().numpy(), Bref) <= tol
    assert rel_err(Sbar.cpu().numpy(), Sref) <= tol
    Sb = Sbar.cpu().numpy()
    assert np.array_equal(Sb, np.transpose(Sb, (0, 2, 1)))          # exactly symmetric

    ybar = np.zeros(N)
    ybar[0] = 0.75
    Lg = AD.logdet_adjoint(F, ybar).cpu().numpy()
    assert rel_err(Lg, Sinv * mpmath.mpf(0.75)) <= tol

    Lbar = lower_of(rand_mf(rng, N, n, n))
    Cg = AD.cholesky_adjoint(F, cu(Lbar)).cpu().numpy()
    Lm = mpmath.cholesky(Sm)
    Linv = Lm ** -1
    P = Lm.T * to_mp(Lbar)
    for i in range(n):
        for j in range(n):
            if j > i:
                P[i, j] = 0
            elif i == j:
                P[i, j] = P[i, j] / 2
    A = Linv.T * P * Linv
    assert rel_err(Cg, (A + A.T) / 2) <= tol


def _fd_dir(fn, S, E, N):
    """(f(S + hE) - f(S - hE)) / 2h in Float64xN on the device."""
    from mfla import api as M

    hE = np.zeros_like(S)
    hE[0] = E * 2.0 ** -H_EXP[N]
    Sp = M.arith("add", cu(S).reshape(N, -1), cu(hE).reshape(N, -1)).reshape(S.shape)
    Sm = M.arith("sub", cu(S).reshape(N, -1), cu(hE).reshape(N, -1)).reshape(S.shape)
    fp, fm = fn(Sp), fn(Sm)
    d = M.arith("sub", fp.reshape(N, -1), fm.reshape(N, -1)).cpu().numpy() * 2.0 ** (H_EXP[N] - 1)   # exact scaling
    return d.reshape(fp.shape)


def _exact_pairing_gap(oracle, pairs, N):
    """Exactly sum_k <P_k, Q_k> (Q may carry a sign), rounded; and its scale."""
    L = max(max(p.shape[0], q.shape[0]) for p, q in pairs)

    def flat(x):
        x = x.reshape(x.shape[0], -1)
        return np.pad(x, ((0, L - x.shape[0]), (0, 0)))

    A = np.concatenate([flat(p) for p, _ in pairs], axis=1)
    Bv = np.concatenate([flat(q) for _, q in pairs], axis=1)
    K = A.shape[1]
    A3 = np.ascontiguousarray(A.reshape(L, K, 1))   # 1 x K row vector, planar [col, row]
    B3 = np.ascontiguousarray(Bv.reshape(L, 1, K))  # K x 1 column vector
    val, scale = oracle.exact_axpb(A3, B3, None, rows=np.array([0]), cols=np.array([0]))
    return float(val[0]), float(scale[0])


def _sym_dir(rng, n):
    E = rng.standard_normal((n, n))
    return (E + E.T) / 2


@pytest.mark.parametrize("N", [3, 4])
def test_solve_rule_matches_finite_differences(oracle, N):
    from mfla import api as M
    from mfla import autodiff as AD

    rng = np.random.default_rng(90 + N)
    n, nrhs = 48, 4
    S = spd(oracle, rng, N, n)
    B = rand_mf(rng, N, n, nrhs)
    W = rand_mf(rng, N, n, nrhs)
    F = M.cholesky_(cu(S))
    X = M.solve(F, cu(B))
    Bbar, Sbar = AD.solve_adjoint(F, X, cu(W))
    E = _sym_dir(rng, n)                             # matrix direction (symmetric)
    fd = _fd_dir(lambda Sx: M.solve(M.cholesky_(Sx.contiguous()), cu(B)), S, E.T.copy(), N)
    gap, scale = _exact_pairing_gap(oracle, [(W, fd), (Sbar.cpu().numpy(), -E.T[None])], N)
    assert abs(gap) <= FD_TOL[N] * scale, (gap, scale)
    EB = rng.standard_normal((nrhs, n))              # right-hand-side direction
    hB = np.zeros_like(B)
    hB[0] = EB * 2.0 ** -H_EXP[N]
    Bp = M.arith("add", cu(B).reshape(N, -1), cu(hB).reshape(N, -1)).reshape(B.shape)
    Bm = M.arith("sub", cu(B).reshape(N, -1), cu(hB).reshape(N, -1)).reshape(B.shape)
    d = M.arith("sub", M.solve(F, Bp).reshape(N, -1), M.solve(F, Bm).reshape(N, -1)).cpu().numpy() * 2.0 ** (H_EXP[N] - 1)
    gap, scale = _exact_pairing_gap(oracle, [(W, d.reshape(B.shape)), (Bbar.cpu().numpy(), -EB[None])], N)
    assert abs(gap) <= FD_TOL[N] * scale, (gap, scale)


@pytest.mark.parametrize("N", [3, 4])
def test_logdet_and_cholesky_rules_match_finite_differences(oracle, N):
    from mfla import api as M
    from mfla import autodiff as AD

    rng = np.random.default_rng(110 + N)
    n = 40
    S = spd(oracle, rng, N, n)
    F = M.cholesky_(cu(S))
    E = _sym_dir(rng, n)
    one = np.zeros(N)
    one[0] = 1.0
    G = AD.logdet_adjoint(F, one).cpu().numpy()
    fd = _fd_dir(lambda Sx: torch.from_numpy(AD.logdet(M.cholesky_(Sx.contiguous()))).cuda().reshape(N, 1, 1), S, E.T.copy(), N)
    gap, scale = _exact_pairing_gap(oracle, [(fd, np.ones((1, 1, 1))), (G, -E.T[None])], N)
    assert abs(gap) <= FD_TOL[N] * scale, (gap, scale)

    Lbar = lower_of(rand_mf(rng, N, n, n))
    C = AD.cholesky_adjoint(F, cu(Lbar)).cpu().numpy()
    fd = _fd_dir(lambda Sx: cu(lower_of(M.cholesky_(Sx.contiguous()).factors.cpu().numpy())), S, E.T.copy(), N)
    gap, scale = _exact_pairing_gap(oracle, [(Lbar, fd), (C, -E.T[None])], N)
    assert abs(gap) <= FD_TOL[N] * scale, (gap, scale)


def test_autograd_functions_call_the_adjoint_kernels(oracle):
    from mfla import api as M
    from mfla import autodiff as AD

    N, n, nrhs = 4, 32, 3
    rng = np.random.default_rng(5)
    S = cu(spd(oracle, rng, N, n)).requires_grad_(True)
    B = cu(rand_mf(rng, N, n, nrhs)).requires_grad_(True)
    W = cu(rand_mf(rng, N, n, nrhs))
    X = AD.cholesky_solve(S, B)
    X.backward(W)
    F = M.cholesky_(S.detach().clone())
    Bb, Sb = AD.solve_adjoint(F, M.solve(F, B.detach()), W)
    assert torch.equal(B.grad, Bb) and torch.equal(S.grad, Sb)

    S2 = S.detach().clone().requires_grad_(True)
    y = AD.logdet_fn(S2)
    assert torch.equal(y.cpu(), torch.from_numpy(AD.logdet(F)))
    yb = torch.zeros(N, dtype=torch.float64, device="cuda")
    yb[0] = 2.0
    y.backward(yb)
    assert torch.equal(S2.grad, AD.logdet_adjoint(F, yb.cpu().numpy()))

    S3 = S.detach().clone().requires_grad_(True)
    Lt = AD.cholesky_fn(S3)
    Lb = cu(lower_of(rand_mf(rng, N, n, n)))
    Lt.backward(Lb)
    assert torch.equal(S3.grad, AD.cholesky_adjoint(F, Lb))


def test_adjoint_argument_errors(oracle):
    from mfla import api as M
    from mfla import autodiff as AD

    N, n = 2, 8
    rng = np.random.default_rng(1)
    F = M.cholesky_(cu(spd(oracle, rng, N, n)))
    X = cu(rand_mf(rng, N, n, 2))
    with pytest.raises(M.DimensionMismatch):
        AD.solve_adjoint(F, X, cu(rand_mf(rng, N, n, 3)))
    bad = M.MFCholesky(F.factors, 3)
    with pytest.raises(M.PosDefException):
        AD.logdet(bad)
