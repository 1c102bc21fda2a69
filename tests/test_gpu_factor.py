"""GPU parity of the factorizations, solves and residual path.

* Cholesky / LU / LDLT on seeded matrices: factors, pivots, blocks, D and
  status bitwise equal to the oracle (same per-element order and networks),
  across limb counts, block sizes and sizes that are not block multiples.
* Solves bitwise equal to the oracle in MFLA's exact substitution order.
* The reference's known-answer tests (tests/kat_cases.py) on the GPU.
* Residual, mixed residual and backward error bitwise equal to the oracle.
"""
import numpy as np
import pytest
import torch

from backends import GpuBackend
from kat_cases import run_all
from mfutil import assert_bitwise, rand_mf, renorm

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.fixture(scope="module")
def M():
    import mfla

    assert torch.cuda.is_available()
    return mfla


def spd(oracle, rng, N, n):
    """S = R R^T + n I formed in Float64xN by the oracle (full symmetric)."""
    R = rand_mf(rng, N, n, n)
    S = np.zeros((N, n, n))
    oracle.gemm(S, R, np.ascontiguousarray(np.transpose(R, (0, 2, 1))), 1.0, 0.0, fused=False)
    shift = np.zeros((N, n))
    shift[0] = float(n)
    diag = np.ascontiguousarray(S[:, np.arange(n), np.arange(n)])
    S[:, np.arange(n), np.arange(n)] = oracle.arith("add", diag, shift)
    return S


def kkt(rng, N, m):
    p = m // 2
    H = rand_mf(rng, N, m, m)
    Hs = H + np.transpose(H, (0, 2, 1))
    Hs[0, np.arange(m), np.arange(m)] += 2.0 * m
    A = rand_mf(rng, N, p, m)
    n = m + p
    K = np.zeros((N, n, n))
    K[:, :m, :m] = Hs
    K[:, :m, m:] = A                            # rows m.., columns 0..m-1: A
    K[:, m:, :m] = np.transpose(A, (0, 2, 1))   # rows 0..m-1, columns m..: A^T
    K[0, np.arange(m, n), np.arange(m, n)] = -1e-8
    return K


@pytest.mark.parametrize("N", [2, 4])
@pytest.mark.parametrize("bad", [5, 150, 290])
def test_cholesky_failure_in_a_late_panel_bitwise(M, oracle, N, bad):
    """A breakdown in panel j leaves the matrix exactly as MFLA's sequential
    order does: every trailing update of panels < j applied, nothing after
    (the look-ahead's overlapped updates must neither skip nor add work)."""
    rng = np.random.default_rng(bad + N)
    n, block = 300, 16
    S = spd(oracle, rng, N, n)
    S[:, bad, bad] = 0.0
    S[0, bad, bad] = -1.0                       # not positive definite at column `bad`
    ref = S.copy()
    info_ref = oracle.potrf(ref, block)
    Sd = dev(S)
    F = M.cholesky_(Sd, check=False, config=M.KernelConfig(cholesky_block=block))
    assert F.info == info_ref and 0 < info_ref <= bad + 1
    assert_bitwise(host(Sd), ref, f"failed cholesky x{N} at {info_ref}")


@pytest.mark.parametrize("N", [2, 3, 4])
def test_reference_kats_on_gpu(M, N):
    res = run_all(GpuBackend(), N)
    failed = [m for ok, m in res if not ok]
    assert not failed, failed


@pytest.mark.parametrize("N", [1, 2, 3, 4])
@pytest.mark.parametrize("n,block", [(37, 16), (64, 16), (100, 7), (130, 32), (300, 16)])
def test_cholesky_bitwise(M, oracle, N, n, block):
    rng = np.random.default_rng(n * 10 + N + block)
    S = spd(oracle, rng, N, n)
    S[:, np.triu_indices(n, 1)[1], np.triu_indices(n, 1)[0]] = np.nan     # only the lower triangle is read
    ref = S.copy()
    info_ref = oracle.potrf(ref, block)
    Sd = dev(S)
    F = M.cholesky_(Sd, config=M.KernelConfig(cholesky_block=block))
    assert F.info == info_ref == 0
    assert_bitwise(host(Sd), ref, f"cholesky x{N} n={n} b={block}")
    B = rand_mf(rng, N, n, 3)
    Xref = oracle.potrs(ref, B.copy())
    X = M.solve(F, dev(B), config=M.KernelConfig(trsm_order="exact"))
    assert_bitwise(host(X), Xref, "potrs")
    b = rand_mf(rng, N, n)
    x = M.solve(F, dev(b), config=M.KernelConfig(trsm_order="exact"))
    assert_bitwise(host(x), oracle.potrs(ref, b.copy()), "potrs vector")


@pytest.mark.parametrize("N", [2, 4])
@pytest.mark.parametrize("bad", [5, 150, 290])
def test_lu_failure_in_a_late_panel_bitwise(M, oracle, N, bad):
    """A zero pivot in panel j leaves the matrix and the pivots exactly as
    MFLA's order does (whole-row interchanges up to the failing column, every
    trailing update of panels < j, nothing after), with look-ahead active."""
    rng = np.random.default_rng(bad * 3 + N)
    n, block = 300, 16
    A = rand_mf(rng, N, n, n)
    A[:, bad, :] = 0.0                           # column `bad` stays exactly zero
    ref = A.copy()
    info_ref, ipiv_ref, _ = oracle.getrf(ref, block)
    Ad = dev(A)
    F = M.lu_(Ad, check=False, config=M.KernelConfig(lu_block=block))
    assert F.info == info_ref == bad + 1
    assert np.array_equal(F.ipiv.cpu().numpy(), ipiv_ref)
    assert_bitwise(host(Ad), ref, f"failed lu x{N} at {info_ref}")


@pytest.mark.parametrize("N", [1, 2, 3, 4])
@pytest.mark.parametrize("m,n,block", [(40, 40, 16), (67, 67, 16), (90, 50, 8), (50, 90, 16), (129, 129, 32), (300, 300, 16),
                                       (300, 200, 16), (200, 300, 16), (257, 257, 8)])
def test_lu_bitwise(M, oracle, N, m, n, block):
    rng = np.random.default_rng(m * 100 + n + N + block)
    A = rand_mf(rng, N, m, n)
    ref = A.copy()
    info_ref, ipiv_ref, om_ref = oracle.getrf(ref, block)
    Ad = dev(A)
    F = M.lu_(Ad, config=M.KernelConfig(lu_block=block))
    assert F.info == info_ref == 0
    assert np.array_equal(F.ipiv.cpu().numpy(), ipiv_ref)
    assert_bitwise(F.original_maximum, om_ref, "original maximum")
    assert_bitwise(host(Ad), ref, f"lu x{N} {m}x{n} b={block}")
    if m == n:
        B = rand_mf(rng, N, n, 4)
        X = M.solve(F, dev(B), config=M.KernelConfig(trsm_order="exact"))
        assert_bitwise(host(X), oracle.getrs(ref, ipiv_ref, B.copy()), "getrs")


@pytest.mark.parametrize("N", [2, 3, 4])
@pytest.mark.parametrize("m,strategy,block", [(20, "unblocked", 0), (60, "unblocked", 0), (60, "blocked", 0), (90, "blocked", 5)])
def test_ldlt_kkt_bitwise(M, oracle, N, m, strategy, block):
    rng = np.random.default_rng(m + N * 7 + block)
    K = kkt(rng, N, m)
    n = K.shape[1]
    ref = K.copy()
    info_ref, dsub_ref, piv_ref, blk_ref, om_ref = oracle.sytrf(ref, strategy, block, 1)
    Kd = dev(K)
    F = M.ldlt_(Kd, config=M.KernelConfig(ldlt_strategy=strategy, ldlt_block=block, ldlt_blocked_crossover=1))
    assert F.info == info_ref == 0
    assert np.array_equal(F.pivots.cpu().numpy(), piv_ref), "pivots"
    assert np.array_equal(F.blocks.cpu().numpy(), blk_ref), "blocks"
    assert_bitwise(host(F.dsub), dsub_ref, "dsub")
    assert_bitwise(host(Kd), ref, f"ldlt x{N} n={n} {strategy}")
    B = rand_mf(rng, N, n, 4)
    X = M.solve(F, dev(B), config=M.KernelConfig(trsm_order="exact"))
    assert_bitwise(host(X), oracle.sytrs(ref, dsub_ref, piv_ref, blk_ref, B.copy()), "sytrs")
    inertia = M.factor_inertia(F)
    assert inertia["positive"] == m and inertia["negative"] == n - m and inertia["zero"] == 0


@pytest.mark.parametrize("N", [2, 4])
@pytest.mark.parametrize("zero", [0, 17, 44])
def test_ldlt_failure_state_bitwise(M, oracle, N, zero):
    """A zero row / column stops the unblocked Bunch-Kaufman factorization
    when it becomes the pivot column: the returned matrix (both triangles),
    pivots, blocks and dsub are MFLA's at that point."""
    rng = np.random.default_rng(zero + 3 * N)
    K = kkt(rng, N, 30)
    K[:, zero, :] = 0.0
    K[:, :, zero] = 0.0
    ref = K.copy()
    info_ref, dsub_ref, piv_ref, blk_ref, _ = oracle.sytrf(ref, "unblocked", 0, 1)
    assert info_ref > 0
    Kd = dev(K)
    F = M.ldlt_(Kd, check=False, config=M.KernelConfig(ldlt_strategy="unblocked"))
    assert F.info == info_ref
    assert np.array_equal(F.pivots.cpu().numpy(), piv_ref) and np.array_equal(F.blocks.cpu().numpy(), blk_ref)
    assert_bitwise(host(F.dsub), dsub_ref, "dsub")
    assert_bitwise(host(Kd), ref, f"failed ldlt x{N} zero column {zero}")


@pytest.mark.parametrize("N", [1, 2, 3, 4])
@pytest.mark.parametrize("n,block", [(150, 0), (260, 3), (200, 70), (1100, 0)])
def test_ldlt_blocked_indefinite_bitwise(M, oracle, N, n, block):
    """Blocked lazy Bunch-Kaufman on a random symmetric indefinite matrix:
    interchanges and 2x2 pivots in most panels. Block 70 takes the panel's
    per-element Schur entries (more panel columns than its LDS factor slots);
    the others take the per-column factors. n = 1100 gives each of the
    panel's 512 threads up to three rows of a column."""
    rng = np.random.default_rng(n + 11 * N + block)
    H = rand_mf(rng, N, n, n)
    K = renorm(H + np.transpose(H, (0, 2, 1)))
    ref = K.copy()
    info_ref, dsub_ref, piv_ref, blk_ref, om_ref = oracle.sytrf(ref, "blocked", block, 1, threads=8)
    assert info_ref == 0 and (blk_ref == 2).any() and (piv_ref != np.arange(1, n + 1)).any()
    Kd = dev(K)
    F = M.ldlt_(Kd, config=M.KernelConfig(ldlt_strategy="blocked", ldlt_block=block, ldlt_blocked_crossover=1))
    assert F.info == 0
    assert np.array_equal(F.pivots.cpu().numpy(), piv_ref), "pivots"
    assert np.array_equal(F.blocks.cpu().numpy(), blk_ref), "blocks"
    assert_bitwise(host(F.dsub), dsub_ref, "dsub")
    assert_bitwise(host(Kd), ref, f"blocked ldlt x{N} n={n} b={block}")


@pytest.mark.parametrize("N", [2, 3, 4])
def test_ldlt_unblocked_indefinite_many_rows_bitwise(M, oracle, N):
    """The right-looking route (forced) on a random symmetric indefinite
    matrix with two rows of a column per thread of its 512."""
    n = 700
    rng = np.random.default_rng(17 + N)
    H = rand_mf(rng, N, n, n)
    K = renorm(H + np.transpose(H, (0, 2, 1)))
    ref = K.copy()
    info_ref, dsub_ref, piv_ref, blk_ref, _ = oracle.sytrf(ref, "unblocked", 0, 1, threads=8)
    assert info_ref == 0 and (blk_ref == 2).any()
    Kd = dev(K)
    F = M.ldlt_(Kd, config=M.KernelConfig(ldlt_strategy="unblocked"))
    assert F.info == 0
    assert np.array_equal(F.pivots.cpu().numpy(), piv_ref) and np.array_equal(F.blocks.cpu().numpy(), blk_ref)
    assert_bitwise(host(F.dsub), dsub_ref, "dsub")
    assert_bitwise(host(Kd), ref, f"unblocked ldlt x{N} n={n}")


@pytest.mark.parametrize("N", [2, 4])
def test_ldlt_blocked_spd_many_rows_bitwise(M, oracle, N):
    """Blocked LDLT of a diagonally dominant matrix (1x1 pivots, no
    interchanges: the panel reuses its column-pass values) with up to three
    rows of each column per panel thread."""
    n = 1300
    rng = np.random.default_rng(91 + N)
    H = rand_mf(rng, N, n, n)
    K = renorm(H + np.transpose(H, (0, 2, 1)))
    K[0, np.arange(n), np.arange(n)] += 4.0 * n
    ref = K.copy()
    info_ref, dsub_ref, piv_ref, blk_ref, _ = oracle.sytrf(ref, "blocked", 0, 1, threads=8)
    assert info_ref == 0 and (blk_ref == 1).all() and (piv_ref == np.arange(1, n + 1)).all()
    Kd = dev(K)
    F = M.ldlt_(Kd, config=M.KernelConfig(ldlt_strategy="blocked", ldlt_blocked_crossover=1))
    assert F.info == 0
    assert np.array_equal(F.pivots.cpu().numpy(), piv_ref) and np.array_equal(F.blocks.cpu().numpy(), blk_ref)
    assert_bitwise(host(Kd), ref, f"blocked ldlt x{N} n={n} diagonally dominant")


@pytest.mark.parametrize("N", [2, 4])
@pytest.mark.parametrize("zero", [5, 37, 90])
def test_ldlt_blocked_failure_state_bitwise(M, oracle, N, zero):
    """A zero row / column stops the blocked factorization inside the panel
    that reaches it: info, pivots, blocks, dsub and the matrix are MFLA's."""
    rng = np.random.default_rng(zero + 5 * N)
    K = kkt(rng, N, 70)
    K[:, zero, :] = 0.0
    K[:, :, zero] = 0.0
    ref = K.copy()
    info_ref, dsub_ref, piv_ref, blk_ref, _ = oracle.sytrf(ref, "blocked", 8, 1)
    assert info_ref > 0
    Kd = dev(K)
    F = M.ldlt_(Kd, check=False, config=M.KernelConfig(ldlt_strategy="blocked", ldlt_block=8, ldlt_blocked_crossover=1))
    assert F.info == info_ref
    assert np.array_equal(F.pivots.cpu().numpy(), piv_ref) and np.array_equal(F.blocks.cpu().numpy(), blk_ref)
    assert_bitwise(host(F.dsub), dsub_ref, "dsub")
    assert_bitwise(host(Kd), ref, f"failed blocked ldlt x{N} zero column {zero}")


def _ldlt_with_panel_workgroups(M, K, cap, **cfg):
    """ldlt! with the panel's workgroup cap set through the environment (0:
    the 512-thread single-workgroup panel kernel)."""
    import os

    old = os.environ.get("MFLA_LDLT_PANEL_WORKGROUPS")
    os.environ["MFLA_LDLT_PANEL_WORKGROUPS"] = str(cap)
    try:
        Kd = dev(K)
        F = M.ldlt_(Kd, check=False, config=M.KernelConfig(ldlt_strategy="blocked", ldlt_blocked_crossover=1, **cfg))
        torch.cuda.synchronize()
    finally:
        if old is None:
            del os.environ["MFLA_LDLT_PANEL_WORKGROUPS"]
        else:
            os.environ["MFLA_LDLT_PANEL_WORKGROUPS"] = old
    return F, host(Kd)


def _ldlt_with_env(M, K, env, workspace=None, **cfg):
    """ldlt! (blocked route) with environment variables set for the call."""
    import os

    old = {k: os.environ.get(k) for k in env}
    os.environ.update({k: str(v) for k, v in env.items()})
    try:
        Kd = dev(K)
        F = M.ldlt_(Kd, check=False, config=M.KernelConfig(ldlt_strategy="blocked", ldlt_blocked_crossover=1, **cfg),
                    workspace=workspace)
        torch.cuda.synchronize()
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v
    return F, host(Kd)


def _assert_same_ldlt(Fa, Ka, Fb, Kb, what):
    assert Fa.info == Fb.info, what
    assert np.array_equal(Fa.pivots.cpu().numpy(), Fb.pivots.cpu().numpy()), f"{what}: pivots"
    assert np.array_equal(Fa.blocks.cpu().numpy(), Fb.blocks.cpu().numpy()), f"{what}: blocks"
    assert_bitwise(host(Fa.dsub), host(Fb.dsub), f"{what}: dsub")
    assert_bitwise(Ka, Kb, f"{what}: matrix")


@pytest.mark.parametrize("N", [1, 2, 3, 4])
@pytest.mark.parametrize("kind,n,block", [("indefinite", 900, 0), ("indefinite", 700, 40), ("dominant", 1000, 0),
                                          ("indefinite", 2600, 0)])
def test_ldlt_panel_over_workgroups_equals_one_workgroup(M, N, kind, n, block):
    """The multi-workgroup panel kernel (rows below the panel spread over up
    to ceil(rows / 256) workgroups, grid barriers between the phases of a
    step; one workgroup under a cap of 1, several rows per thread under a cap
    of 2) against the 512-thread single-workgroup kernel, bitwise:
    interchanges and 2x2 pivots (indefinite), 1x1 steps without interchange
    (dominant), the per-element entries of a panel wider than its LDS slots
    (block 40), and 11 workgroups (n = 2600)."""
    if n > 1000 and N != 2:
        pytest.skip("the 11-workgroup case runs for two limbs")
    rng = np.random.default_rng(n + 13 * N + block)
    H = rand_mf(rng, N, n, n)
    K = renorm(H + np.transpose(H, (0, 2, 1)))
    if kind == "dominant":
        K[0, np.arange(n), np.arange(n)] += 4.0 * n
    F0, K0 = _ldlt_with_panel_workgroups(M, K, 0, ldlt_block=block)
    assert F0.info == 0
    if kind == "indefinite":
        assert (F0.blocks.cpu().numpy() == 2).any()
    for cap in (64, 1, 2):
        Fg, Kg = _ldlt_with_panel_workgroups(M, K, cap, ldlt_block=block)
        _assert_same_ldlt(Fg, Kg, F0, K0, f"ldlt x{N} n={n} {kind} b={block} cap {cap}")


@pytest.mark.parametrize("N", [2, 4])
@pytest.mark.parametrize("zero", [150, 301])
def test_ldlt_panel_over_workgroups_failure_state(M, oracle, N, zero):
    """A zero row / column reached inside a panel that runs over several
    workgroups: info, pivots, blocks, dsub and the matrix are MFLA's (the
    oracle) and those of the single-workgroup panel."""
    rng = np.random.default_rng(zero + 7 * N)
    K = kkt(rng, N, 420)
    n = K.shape[1]
    K[:, zero, :] = 0.0
    K[:, :, zero] = 0.0
    ref = K.copy()
    info_ref, dsub_ref, piv_ref, blk_ref, _ = oracle.sytrf(ref, "blocked", 0, 1, threads=8)
    assert info_ref > 0
    Fg, Kg = _ldlt_with_panel_workgroups(M, K, 64)
    F1, K1 = _ldlt_with_panel_workgroups(M, K, 0)
    assert Fg.info == info_ref
    assert np.array_equal(Fg.pivots.cpu().numpy(), piv_ref) and np.array_equal(Fg.blocks.cpu().numpy(), blk_ref)
    assert_bitwise(host(Fg.dsub), dsub_ref, "dsub")
    assert_bitwise(Kg, ref, f"failed blocked ldlt x{N} n={n} zero column {zero}")
    _assert_same_ldlt(Fg, Kg, F1, K1, f"failed ldlt x{N} zero {zero}")


def _lookahead_case(rng, N, kind, n):
    H = rand_mf(rng, N, n, n)
    K = renorm(H + np.transpose(H, (0, 2, 1)))
    idx = np.arange(n)
    if kind == "dominant":
        K[0, idx, idx] += 4.0 * n
    elif kind == "dominant_then_indefinite":
        # 1x1 steps without interchange up to column 403 (inside a panel),
        # then steps that need the row-max pass
        K[0, idx[:403], idx[:403]] += 4.0 * n
    elif kind == "dominant_zero":
        K[0, idx, idx] += 4.0 * n
        K[:, 301, :] = 0.0
        K[:, :, 301] = 0.0
    return K


@pytest.mark.parametrize("N", [1, 2, 3, 4])
@pytest.mark.parametrize("kind,n", [("dominant", 1000), ("indefinite", 900), ("dominant_then_indefinite", 800),
                                    ("dominant_zero", 700)])
def test_ldlt_lookahead_equals_serial(M, N, kind, n):
    """Look-ahead (the panels on a few CUs, each trailing update split into
    the next panel's columns and the rest, which runs beside the next panel)
    against the serial order, bitwise: throughout (dominant), stopped panels
    restored and run again (a panel that reaches a column needing the row-max
    pass while the previous update runs: at a panel's first column
    (indefinite) or inside it after 1x1 steps (column 403)), and a zero
    column met beside a running update (info 302)."""
    rng = np.random.default_rng(n + 31 * N)
    K = _lookahead_case(rng, N, kind, n)
    F0, K0 = _ldlt_with_env(M, K, {"MFLA_LDLT_LOOKAHEAD": 0})
    if kind == "dominant_zero":
        assert F0.info == 302
    else:
        assert F0.info == 0
    if kind != "dominant":
        assert (F0.blocks.cpu().numpy() == 2).any() or kind == "dominant_zero"
    for env in ({"MFLA_LDLT_LOOKAHEAD": 1}, {"MFLA_LDLT_LOOKAHEAD": 1, "MFLA_LDLT_PANEL_WORKGROUPS": 2}):
        F1, K1 = _ldlt_with_env(M, K, env)
        _assert_same_ldlt(F1, K1, F0, K0, f"ldlt x{N} n={n} {kind} {env}")
    # the weighted panels' three rotating parts taken from a workspace, twice
    ws = M.MFWorkspace(N, factor_capacity=n, ldlt_block_capacity=M.ldlt_plan(N, n).block_size)
    for rep in range(2):
        F2, K2 = _ldlt_with_env(M, K, {"MFLA_LDLT_LOOKAHEAD": 1}, workspace=ws)
        _assert_same_ldlt(F2, K2, F0, K0, f"ldlt x{N} n={n} {kind} workspace call {rep}")


def test_ldlt_lookahead_against_oracle(M, oracle):
    """The look-ahead route against the oracle (MFLA's blocked ldlt!) on a
    matrix whose panels stop inside a panel and at panel starts."""
    rng = np.random.default_rng(77)
    K = _lookahead_case(rng, 2, "dominant_then_indefinite", 600)
    ref = K.copy()
    info_ref, dsub_ref, piv_ref, blk_ref, _ = oracle.sytrf(ref, "blocked", 0, 1, threads=8)
    F, Kg = _ldlt_with_env(M, K, {"MFLA_LDLT_LOOKAHEAD": 1})
    assert F.info == info_ref == 0
    assert np.array_equal(F.pivots.cpu().numpy(), piv_ref) and np.array_equal(F.blocks.cpu().numpy(), blk_ref)
    assert_bitwise(host(F.dsub), dsub_ref, "dsub")
    assert_bitwise(Kg, ref, "look-ahead ldlt x2 n=600")


def test_status_codes_and_check(M):
    A = torch.zeros((2, 3, 3), dtype=torch.float64, device="cuda")
    A[0, 1, 2] = float("nan")                       # (row 2, col 1): lower triangle
    F = M.cholesky_(A.clone(), check=False)
    assert F.info == -1 and M.factor_state(F) == "nonfinite_input"
    with pytest.raises(M.DomainError):
        M.cholesky_(A.clone())
    A2 = torch.zeros((2, 3, 3), dtype=torch.float64, device="cuda")
    A2[0, 2, 0] = float("nan")                      # (row 0, col 2): upper -- never read
    A2[0, [0, 1, 2], [0, 1, 2]] = 1.0
    assert M.cholesky_(A2.clone()).info == 0
    L = torch.zeros((2, 3, 3), dtype=torch.float64, device="cuda")
    L[0, 0, 0] = 1.0
    Fl = M.lu_(L, check=False)
    assert Fl.info == 2 and M.factor_state(Fl) == "singular"
    with pytest.raises(M.SingularException):
        M.lu_(L.clone())
    # a failed factor raises before the destination is touched (solve.jl:381-400)
    b = torch.ones((2, 3), dtype=torch.float64, device="cuda")
    dst = torch.full((2, 3), 7.0, dtype=torch.float64, device="cuda")
    with pytest.raises(M.SingularException):
        M.ldiv_(dst, Fl, b)
    assert torch.all(dst == 7.0)


@pytest.mark.parametrize("N", [2, 3, 4])
def test_residual_paths_bitwise(M, oracle, N):
    rng = np.random.default_rng(55 + N)
    m, n = 41, 29
    A, x, b = rand_mf(rng, N, m, n), rand_mf(rng, N, n), rand_mf(rng, N, m)
    r = torch.full((N, m), float("nan"), dtype=torch.float64, device="cuda")
    M.residual_(r, dev(A), dev(x), dev(b))
    assert_bitwise(host(r), oracle.gemv(b.copy(), A, x, -1.0, 1.0), "residual vector (gemv)")
    X, B = rand_mf(rng, N, n, 3), rand_mf(rng, N, m, 3)
    R = torch.full((N, 3, m), float("nan"), dtype=torch.float64, device="cuda")
    M.residual_(R, dev(A), dev(X), dev(B))
    assert_bitwise(host(R), oracle.gemm(B.copy(), A, X, -1.0, 1.0, fused=True), "residual matrix (gemm)")
    S = rand_mf(rng, N, n, n)
    rs = dev(b[:, :n].copy())
    M.residual_(rs, dev(S), dev(x), rs, uplo="lower")          # in place on b
    assert_bitwise(host(rs), oracle.symv(b[:, :n].copy(), S, x, -1.0, 1.0, "lower"), "residual symmetric")
    be = M.normwise_backward_error(dev(A), dev(X), dev(B), R)
    ref = oracle.backward_error(A, X, B, host(R))
    assert_bitwise(be, ref, "backward error")
    if N < 4:
        for NR in range(N + 1, 5):
            Rm = torch.zeros((NR, 3, m), dtype=torch.float64, device="cuda")
            M.residual_mixed_(Rm, dev(A), dev(X), dev(B))
            assert_bitwise(host(Rm), oracle.residual_mixed(np.zeros((NR, 3, m)), A, X, B), f"mixed x{N}->x{NR}")


def test_workspace_reuse_bitwise(M, oracle):
    rng = np.random.default_rng(3)
    ws = M.MFWorkspace(3, factor_capacity=10, ldlt_block_capacity=4)
    K = kkt(rng, 3, 40)
    outs = []
    for w in (None, ws, ws):
        Kd = dev(K)
        F = M.ldlt_(Kd, config=M.KernelConfig(ldlt_strategy="blocked", ldlt_block=6, ldlt_blocked_crossover=1), workspace=w)
        outs.append(host(Kd))
        assert F.info == 0
    M.ensure_workspace_capacity_(ws, factor_capacity=100, ldlt_block_capacity=6)
    assert M.workspace_capacity(ws) == {"factor": 100, "ldlt_block": 6}
    Kd = dev(K)
    M.ldlt_(Kd, config=M.KernelConfig(ldlt_strategy="blocked", ldlt_block=6, ldlt_blocked_crossover=1), workspace=ws)
    outs.append(host(Kd))
    for o in outs[1:]:
        assert_bitwise(o, outs[0], "workspace reuse")


def test_diagnostics_fields(M, oracle):
    rng = np.random.default_rng(8)
    S = spd(oracle, rng, 4, 30)
    F = M.cholesky_(dev(S))
    d = M.factor_diagnostics(F)
    assert d["kind"] == "cholesky" and d["success"] and d["accepted_pivots"] == 30 and d["finite"]
    assert d["state"] == "success" and d["failure_location"] is None
    spread = M.host_arith("div", d["maximum_diagonal"].reshape(4, 1), d["minimum_diagonal"].reshape(4, 1))[:, 0]
    assert np.array_equal(spread, d["diagonal_spread"])
    A = rand_mf(rng, 4, 20, 20)
    A[0, np.arange(20), np.arange(20)] += 4.0
    Fl = M.lu_(dev(A))
    dl = M.factor_diagnostics(Fl)
    info, ipiv, om = oracle.getrf(A.copy(), 16)
    assert np.array_equal(dl["original_maximum"], om) and np.array_equal(dl["pivots"], ipiv)
    assert M.numerical_rank(Fl) == 20
