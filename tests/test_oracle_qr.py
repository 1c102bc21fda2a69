"""CPU checks of the oracle's QR restatement (qr!, rrqr!, apply_q!;
src/factorizations/qr.jl:606-956) before it is trusted as the GPU checker.

* LAPACK's float64 Householder QR uses the same reflector convention
  (beta = -sign(alpha) norm, tau = (beta - alpha) / beta, v scaled by
  1 / (alpha - beta)): the hi limbs of the factor and tau agree with
  dgeqrf / dgeqp3 to float64 accuracy, and the column pivot order agrees
  with dgeqp3 on matrices whose column norms are well separated. A dropped
  term, a wrong sign or a transposed reflector fails these.
* Exact residuals: Q^T Q - I and Q R - A P through the exact long
  accumulator, within the reference's tolerances.
* The reference's known answers (test/runtests.jl:750-1010,
  test/qr_delayed_norm.jl): ranks 4 / 0 / 5 / 4 / 31 / 30 / 40, the
  equal-norm identity permutation, the extreme-scale first pivot, blocked
  ranks equal to the unblocked route's.
"""
import numpy as np
import pytest
import scipy.linalg as sl

EPS = {2: 2.0 ** -104, 3: 2.0 ** -156, 4: 2.0 ** -208}


def mf(A, N):
    """Float64 matrix (m, n) -> planar (N, n, m) with zero lower limbs."""
    out = np.zeros((N, A.shape[1], A.shape[0]))
    out[0] = A.T
    return out


def rdiag(F):
    k = min(F.shape[1], F.shape[2])
    return np.array([F[0, i, i] for i in range(k)])


def rank_of(F, rtol=0.0, atol=0.0):
    """numerical_rank (qr.jl:863-888) on the hi limbs of R's diagonal."""
    d = np.abs(rdiag(F))
    thr = max(atol, rtol * (d.max() if d.size else 0.0))
    r = 0
    for v in d:
        if not v > thr:
            break
        r += 1
    return r


def explicit_q(oracle, F, tau):
    N, _, m = F.shape
    Q = np.zeros((N, m, m))
    Q[0] = np.eye(m)
    oracle.ormqr(F, tau, Q, trans="N")
    return Q


def check_exact(oracle, A, F, tau, perm, work):
    """|Q R - A P| and |Q^T Q - I| via the exact accumulator."""
    N, n, m = A.shape
    Q = explicit_q(oracle, F, tau)
    R = np.zeros_like(A)
    for j in range(n):
        R[:, j, :min(j + 1, m)] = F[:, j, :min(j + 1, m)]
    AP = np.ascontiguousarray(A[:, perm - 1, :])
    rows = np.repeat(np.arange(m), n)
    cols = np.tile(np.arange(n), m)
    val, _ = oracle.exact_axpb(Q, R, AP, sab=1, sc=-1, rows=rows, cols=cols)
    tol = 4096 * work * EPS[N] * max(1.0, np.abs(A[0]).max())
    assert np.abs(val).max() <= tol
    I = np.zeros((N, m, m))
    I[0] = np.eye(m)
    rows = np.repeat(np.arange(m), m)
    cols = np.tile(np.arange(m), m)
    val, _ = oracle.exact_axpb(Q, Q, I, sab=1, sc=-1, rows=rows, cols=cols, atrans=True)
    assert np.abs(val).max() <= 4096 * work * EPS[N]


@pytest.mark.parametrize("N", [2, 3, 4])
@pytest.mark.parametrize("m,n", [(30, 20), (20, 30), (17, 17), (1, 5), (6, 1)])
def test_geqrf_matches_lapack_and_is_exact(oracle, N, m, n):
    rng = np.random.default_rng(m * 7 + n + N)
    A0 = rng.standard_normal((m, n))
    A = mf(A0, N)
    F = A.copy()
    tau = oracle.geqrf(F)
    (raw, tau_l), _ = sl.qr(A0, mode="raw")
    # dlarfg returns tau = 0 for an empty / zero sub-column; MFLA's reflector
    # (qr.jl:606-622) still reflects (tau = 2, beta = -alpha), negating that
    # row of R. Undo that sign before comparing.
    got = F[0].T.copy()
    flip = np.flatnonzero((tau_l == 0) & (np.abs(tau[0] - 2.0) < 1e-12))
    for s_ in flip:
        got[s_, s_:] *= -1.0
    assert np.abs(got - raw).max() <= 1e-13 * max(1.0, np.abs(raw).max())
    keep = np.setdiff1d(np.arange(len(tau_l)), flip)
    assert np.abs(tau[0][keep] - tau_l[keep]).max(initial=0.0) <= 1e-13
    check_exact(oracle, A, F, tau, np.arange(1, n + 1), 64 * max(m, n))


@pytest.mark.parametrize("N", [2, 3, 4])
@pytest.mark.parametrize("m,n,route", [(15, 7, "auto"), (7, 15, "auto"), (9, 9, "auto"), (300, 120, "blocked"),
                                       (300, 120, "unblocked"), (64, 256, "blocked")])
def test_geqp3_matches_lapack_pivots(oracle, N, m, n, route):
    rng = np.random.default_rng(m + 3 * n + N)
    A0 = rng.standard_normal((m, n)) * np.exp(rng.uniform(-3, 3, n))
    A = mf(A0, N)
    F = A.copy()
    info, tau, perm = oracle.geqp3(F, route)
    assert info == 0
    _, R, piv = sl.qr(A0, pivoting=True)
    assert np.array_equal(perm - 1, piv)
    k = min(m, n)
    assert np.allclose(np.abs(rdiag(F)), np.abs(np.diag(R)[:k]), rtol=1e-12, atol=0)
    if m * n <= 4000:
        check_exact(oracle, A, F, tau, perm, 64 * max(m, n))


@pytest.mark.parametrize("N", [2, 3, 4])
def test_rrqr_reference_known_answers(oracle, N):
    rng = np.random.default_rng(N)
    eps = EPS[N]
    # rank-deficient: [R(13x4) R[:,2] 0] -> rank 4 at rtol sqrt(eps)     (runtests.jl:797-808)
    src = rng.standard_normal((13, 4))
    D = np.hstack([src, src[:, 1:2], np.zeros((13, 1))])
    F = mf(D, N)
    info, _, _ = oracle.geqp3(F)
    assert info == 0 and rank_of(F, rtol=np.sqrt(eps)) == 4
    F = mf(np.zeros((5, 8)), N)
    oracle.geqp3(F)
    assert rank_of(F) == 0
    # a perturbation of sqrt(eps)/4 stays visible to the exact test       (:813-818)
    near = rng.standard_normal((12, 5))
    near[:, 4] = near[:, 0]
    A = mf(near, N)
    s4 = oracle.arith("add", A[:, 4, 0:1], np.array([[np.sqrt(eps) / 4]] + [[0.0]] * (N - 1)))
    A[:, 4, 0] = s4[:, 0]
    F = A.copy()
    oracle.geqp3(F)
    assert rank_of(F) == 5 and rank_of(F, rtol=np.sqrt(eps)) == 4
    # equal exact norms keep the smallest original column                  (:822-827)
    E = np.zeros((8, 4))
    E[np.arange(4), np.arange(4)] = 1.0
    _, _, perm = oracle.geqp3(mf(E, N))
    assert list(perm) == [1, 2, 3, 4]
    # extreme column scales: the 1e100 column pivots first                 (:874-883)
    S = rng.standard_normal((10, 4)) * np.array([1e-100, 1e100, 1e-40, 1e40])
    F = mf(S, N)
    _, _, perm = oracle.geqp3(F)
    assert perm[0] == 2 and rank_of(F) == 4


def _blocked_source(rows=500, cols=40):
    r = np.arange(1, rows + 1)[:, None].astype(float)
    c = np.arange(1, cols + 1)[None, :].astype(float)
    return np.sin(r * 0.01 + c * 0.7) + 0.1 * np.cos(r * 0.003 * c)


@pytest.mark.parametrize("N", [2, 3, 4])
def test_blocked_rrqr_reenter_panel_rank_31(oracle, N):
    """runtests.jl:829-871: the blocked DLAQPS route keeps a column that leaves
    and re-enters a panel; rank 31 at rtol 1e-12."""
    A = mf(_blocked_source(), N)
    F = A.copy()
    info, tau, perm = oracle.geqp3(F)                        # 500 x 40: blocked route
    assert info == 0 and sorted(perm) == list(range(1, 41))
    assert rank_of(F, rtol=1e-12) == 31
    T = np.ascontiguousarray(A[:, perm - 1, :])
    oracle.ormqr(F, tau, T, trans="T")
    lead = T[0].T[:40, :]
    assert np.abs(np.tril(lead, -1)).max() <= 4096 * 512 * 500 * EPS[N]


def _adversarial(kind, N, oracle, rows=256, cols=64):
    r = np.arange(1, rows + 1)[:, None].astype(float)
    c = np.arange(1, cols + 1)[None, :].astype(float)
    A = np.zeros((N, cols, rows))
    if kind == "near_tie":
        A[0] = (np.sin(0.017 * r + 0.19 * c) + 0.18 * np.cos(0.011 * r * c) + 0.03 * np.sin(0.003 * r ** 2 + c)).T
    elif kind == "cancellation":
        rr = r[:, 0]
        u = np.sin(0.021 * rr) + 0.4 * np.cos(0.007 * rr)
        v = np.cos(0.013 * rr) - 0.2 * np.sin(0.005 * rr)
        A[0, 0] = u
        uu = np.zeros((N, rows))
        uu[0] = u
        vv = np.zeros((N, rows))
        vv[0] = np.ldexp(1.0, -60) * v
        A[:, 1, :] = oracle.arith("add", uu, vv)
        A[0, 2:] = (np.sin(0.019 * r + 0.17 * c) + 0.11 * np.cos(0.004 * r * c)).T[2:]
    elif kind == "rank_deficient":
        cc = np.arange(1, 31)[None, :].astype(float)
        core = np.sin(0.013 * r + 0.23 * cc) + 0.21 * np.cos(0.005 * r * cc)
        A[0, :30] = core.T
        A[0, 30:40] = core.T[1:11]
    else:  # extreme_scale
        base = np.sin(0.015 * r + 0.07 * c) + 0.17 * np.cos(0.009 * r * c)
        ex = np.array([-300, -150, 0, 150, 300])[(np.arange(cols)) % 5]
        A[0] = (base * np.ldexp(1.0, ex)[None, :]).T
    return A


@pytest.mark.parametrize("N", [2, 4])
@pytest.mark.parametrize("kind", ["near_tie", "cancellation", "rank_deficient", "extreme_scale"])
def test_blocked_rrqr_adversarial(oracle, N, kind):
    """runtests.jl:931-1010: blocked ranks equal the unblocked route's, the
    rank-deficient panel has rank 30, Q^T A P is triangular."""
    A = _adversarial(kind, N, oracle)
    Fb, Fu = A.copy(), A.copy()
    _, tau_b, perm_b = oracle.geqp3(Fb, "blocked")
    _, _, perm_u = oracle.geqp3(Fu, "unblocked")
    rtol = np.sqrt(EPS[N]) if kind == "rank_deficient" else 0.0
    assert rank_of(Fb, rtol=rtol) == rank_of(Fu, rtol=rtol)
    if kind == "rank_deficient":
        assert rank_of(Fb, rtol=rtol) == 30
    T = np.ascontiguousarray(A[:, perm_b - 1, :])
    oracle.ormqr(Fb, tau_b, T, trans="T")
    lead = T[0].T[:64, :]
    scale = max(1.0, np.abs(A[0]).max())
    assert np.abs(np.tril(lead, -1)).max() / scale <= 8192 * 256 * EPS[N]


def test_delayed_norm_panel_rank_40_and_30(oracle):
    """test/qr_delayed_norm.jl:48-93 (Float64x4)."""
    N = 4
    F = mf(_blocked_source(), N)
    oracle.geqp3(F)
    assert rank_of(F, rtol=1e-30) == 40
    r = np.arange(1, 257)[:, None].astype(float)
    cc = np.arange(1, 31)[None, :].astype(float)
    core = np.sin(0.013 * r + 0.23 * cc) + 0.21 * np.cos(0.005 * r * cc)
    D = np.zeros((256, 64))
    D[:, :30] = core
    D[:, 30:40] = core[:, 1:11]
    F = mf(D, N)
    oracle.geqp3(F)
    assert rank_of(F, rtol=np.sqrt(EPS[N])) == 30


def test_nonfinite_input_reports_minus_one(oracle):
    A = mf(np.random.default_rng(0).standard_normal((7, 4)), 3)
    A[0, 2, 1] = np.nan
    F = A.copy()
    info, tau, perm = oracle.geqp3(F)
    assert info == -1 and list(perm) == [1, 2, 3, 4] and np.all(tau == 0)
    assert np.array_equal(np.isnan(F), np.isnan(A)) and np.array_equal(F[~np.isnan(F)], A[~np.isnan(A)])


@pytest.mark.parametrize("N", [2, 4])
def test_apply_q_roundtrip(oracle, N):
    rng = np.random.default_rng(21)
    A = mf(rng.standard_normal((14, 6)), N)
    F = A.copy()
    _, tau, _ = oracle.geqp3(F)
    B = mf(rng.standard_normal((14, 3)), N)
    X = B.copy()
    oracle.ormqr(F, tau, X, trans="T")
    oracle.ormqr(F, tau, X, trans="N")
    d = oracle.arith("sub", X.reshape(N, -1), B.reshape(N, -1))
    assert np.abs(d[0]).max() <= 4096 * 64 * 14 * EPS[N] * np.abs(B[0]).max()
