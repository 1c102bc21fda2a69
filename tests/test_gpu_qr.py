"""qr!, rrqr! (both routes), apply_q!, solve_r! on the GPU: compact factor,
tau and permutation bitwise equal to the oracle (MFLA's per-element order,
qr.jl:87-956), the reference's rank / permutation known answers, and MFLA's
errors."""
import numpy as np
import pytest
import torch

from mfutil import assert_bitwise, rand_mf
from test_oracle_qr import EPS, _adversarial, _blocked_source, mf, rank_of

pytestmark = pytest.mark.gpu


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.cpu().numpy()


@pytest.mark.parametrize("N", [1, 2, 3, 4])
@pytest.mark.parametrize("m,n", [(30, 20), (20, 30), (17, 17), (130, 70), (1, 5), (6, 1), (200, 3)])
def test_qr_bitwise(oracle, N, m, n):
    import mfla as M

    rng = np.random.default_rng(11 * m + n + N)
    A = rand_mf(rng, N, m, n)
    F = A.copy()
    tau = oracle.geqrf(F)
    Ag = cu(A)
    Fq = M.qr_(Ag)
    assert Fq.info == 0 and Fq.kind == "qr"
    assert_bitwise(host(Ag), F, f"qr! factor x{N} {m}x{n}")
    assert_bitwise(host(Fq.tau), tau, f"qr! tau x{N} {m}x{n}")
    assert list(host(Fq.permutation)) == list(range(1, n + 1))


@pytest.mark.parametrize("N", [2, 3, 4])
@pytest.mark.parametrize("m,n,route", [(15, 7, "auto"), (7, 15, "auto"), (9, 9, "auto"), (64, 40, "auto"), (40, 64, "unblocked"),
                                       (300, 120, "auto"), (120, 300, "blocked"), (33, 17, "blocked"),
                                       (120, 700, "blocked"), (60, 1300, "unblocked")])
def test_rrqr_bitwise(oracle, N, m, n, route):
    """Wide cases give each pivot-search thread several columns (700 over
    256 threads on the delayed-norm route, 1300 over 1024 on the hybrid)."""
    import mfla as M

    rng = np.random.default_rng(5 * m + n + N)
    A = rand_mf(rng, N, m, n)
    A *= np.exp(rng.uniform(-3, 3, n))[None, :, None]
    F = A.copy()
    info, tau, perm = oracle.geqp3(F, route)
    Ag = cu(A)
    Fq = M.rrqr_(Ag, route=route)
    assert Fq.info == info == 0
    assert np.array_equal(host(Fq.permutation), perm)
    assert_bitwise(host(Ag), F, f"rrqr! factor x{N} {m}x{n} {route}")
    assert_bitwise(host(Fq.tau), tau, f"rrqr! tau x{N} {m}x{n} {route}")


@pytest.mark.parametrize("N", [2, 4])
@pytest.mark.parametrize("kind", ["near_tie", "cancellation", "rank_deficient", "extreme_scale"])
def test_rrqr_adversarial_bitwise(oracle, N, kind):
    """256 x 64 panels of runtests.jl:149-200 (the blocked route; the
    cancellation case ends a panel early on an unreliable norm downdate)."""
    import mfla as M

    A = _adversarial(kind, N, oracle)
    F = A.copy()
    _, tau, perm = oracle.geqp3(F)
    Ag = cu(A)
    Fq = M.rrqr_(Ag)
    assert np.array_equal(host(Fq.permutation), perm)
    assert_bitwise(host(Ag), F, f"rrqr! {kind} x{N}")
    assert_bitwise(host(Fq.tau), tau, f"rrqr! tau {kind} x{N}")
    if kind == "rank_deficient":
        assert M.numerical_rank(Fq, rtol=np.sqrt(EPS[N])) == 30


@pytest.mark.parametrize("N", [2, 3, 4])
def test_rrqr_reenter_panel_rank_31(oracle, N):
    import mfla as M

    A = mf(_blocked_source(), N)
    Ag = cu(A)
    Fq = M.rrqr_(Ag)
    F = A.copy()
    _, tau, perm = oracle.geqp3(F)
    assert np.array_equal(host(Fq.permutation), perm)
    assert_bitwise(host(Ag), F, f"rrqr! 500x40 x{N}")
    assert M.numerical_rank(Fq, rtol=1e-12) == 31 == rank_of(F, rtol=1e-12)


@pytest.mark.parametrize("N", [2, 3, 4])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_apply_q_bitwise(oracle, N, trans):
    import mfla as M

    rng = np.random.default_rng(7 + N)
    m, n = 70, 30
    A = rand_mf(rng, N, m, n)
    Ag = cu(A)
    Fq = M.rrqr_(Ag)
    B = rand_mf(rng, N, m, 5)
    ref = oracle.ormqr(host(Ag), host(Fq.tau), B.copy(), trans=trans)
    Bg = cu(B)
    M.apply_q_(Bg, Fq, trans=trans)
    assert_bitwise(host(Bg), ref, f"apply_q! {trans} x{N}")
    v = rand_mf(rng, N, m, 1)
    refv = oracle.ormqr(host(Ag), host(Fq.tau), v.copy(), trans=trans)
    vg = cu(v[:, 0, :])
    M.apply_q_(vg, Fq, trans=trans)
    assert_bitwise(host(vg), refv[:, 0, :], f"apply_q! vector {trans} x{N}")


@pytest.mark.parametrize("N", [2, 4])
def test_solve_r_matches_trsm_of_the_leading_block(oracle, N):
    import mfla as M

    rng = np.random.default_rng(30 + N)
    A = rand_mf(rng, N, 14, 6)
    Ag = cu(A)
    Fq = M.rrqr_(Ag)
    R = np.ascontiguousarray(host(Ag)[:, :6, :6])
    for trans in ("N", "T"):
        B = rand_mf(rng, N, 6, 3)
        ref = oracle.trsm(B.copy(), R, 1.0, "left", "upper", trans, "nonunit")
        Bg = cu(B)
        M.solve_r_(Bg, Fq, 6, trans=trans)
        assert_bitwise(host(Bg), ref, f"solve_r! {trans} x{N}")
        x = rand_mf(rng, N, 6, 1)
        refx = oracle.trsm(x.copy(), R, 1.0, "left", "upper", trans, "nonunit")
        xg = cu(x[:, 0, :])
        M.solve_r_(xg, Fq, 6, trans=trans)
        assert_bitwise(host(xg), refx[:, 0, :], f"solve_r! vector {trans} x{N}")


def test_qr_known_answers_and_errors(oracle):
    import mfla as M

    N = 3
    rng = np.random.default_rng(3)
    E = np.zeros((8, 4))
    E[np.arange(4), np.arange(4)] = 1.0
    assert list(M.factor_permutation(M.rrqr_(cu(mf(E, N))))) == [1, 2, 3, 4]
    S = rng.standard_normal((10, 4)) * np.array([1e-100, 1e100, 1e-40, 1e40])
    Fs = M.rrqr_(cu(mf(S, N)))
    assert M.factor_permutation(Fs)[0] == 2 and M.numerical_rank(Fs) == 4
    assert M.numerical_rank(M.rrqr_(cu(mf(np.zeros((5, 8)), N)))) == 0
    d = M.factor_diagnostics(Fs)
    assert d["kind"] == "qr" and d["rank_at_threshold"] == 4 and d["finite"] and d["rdiag"].shape == (N, 4)
    assert M.factor_rdiag(Fs).shape == (N, 4)
    # non-finite input
    bad = mf(rng.standard_normal((7, 4)), N)
    bad[0, 2, 1] = np.nan
    with pytest.raises(M.DomainError):
        M.rrqr_(cu(bad))
    Fb = M.rrqr_(cu(bad), check=False)
    assert Fb.info == -1 and not M.issuccess(Fb) and list(M.factor_permutation(Fb)) == [1, 2, 3, 4]
    with pytest.raises(M.ArgumentError):
        M.numerical_rank(Fb)
    # apply_q! / solve_r! errors (runtests.jl:909-925)
    A = mf(rng.standard_normal((14, 6)), N)
    Fq = M.rrqr_(cu(A))
    with pytest.raises(M.ArgumentError):
        M.apply_q_(torch.zeros((N, 14), dtype=torch.float64, device="cuda"), Fq, trans="bad")
    with pytest.raises(M.DimensionMismatch):
        M.apply_q_(torch.zeros((N, 13), dtype=torch.float64, device="cuda"), Fq)
    with pytest.raises(M.ArgumentError):
        M.apply_q_(Fq.factors[:, 0, :], Fq)                       # aliases the factor
    with pytest.raises(M.ArgumentError):
        M.solve_r_(torch.zeros((N, 6), dtype=torch.float64, device="cuda"), Fq, 7)
    with pytest.raises(M.DimensionMismatch):
        M.solve_r_(torch.zeros((N, 5), dtype=torch.float64, device="cuda"), Fq, 6)
    with pytest.raises(M.ArgumentError):
        M.numerical_rank(Fq, atol=-1.0)
    with pytest.raises(M.ArgumentError):
        M.numerical_rank(Fq, rtol=np.inf)


def test_fixed_permutation_qr(oracle):
    """qr_fixed_permutation.jl: qr!(A, perm) equals qr! of A[:, perm] and
    records the caller's order; bad permutations fail closed."""
    import mfla as M

    N, m, n = 2, 40, 12
    rng = np.random.default_rng(0x5EED)
    A0 = rand_mf(rng, N, m, n)
    perm = list(range(n, 0, -1))
    A1 = cu(A0)
    F1 = M.qr_(A1, perm)
    assert list(M.factor_permutation(F1)) == perm
    A2 = np.ascontiguousarray(A0[:, np.array(perm) - 1, :])
    tau = oracle.geqrf(A2)
    assert_bitwise(host(A1), A2, "qr!(A, perm)")
    assert_bitwise(host(F1.tau), tau, "qr!(A, perm) tau")
    A3 = cu(A0)
    with pytest.raises(M.DimensionMismatch):
        M.qr_(A3, [1] * (n + 1))
    with pytest.raises(M.ArgumentError):
        M.qr_(A3, [1, 1] + list(range(3, n + 1)))
    with pytest.raises(M.ArgumentError):
        M.qr_(A3, [0] + list(range(2, n + 1)))
    assert_bitwise(host(A3), A0, "rejected permutations leave A untouched")
