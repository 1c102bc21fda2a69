"""GPU parity of the BLAS-like kernels against the oracle (bitwise).

Every kernel here reduces each output element in MFLA's ascending order with
the same multi-double network as the oracle, so any difference is a bug.
Shapes include odd sizes, sizes that are not multiples of the tiles, empty
dimensions and the beta == 0 NaN-destination case (runtests.jl:315-396).
"""
import numpy as np
import pytest
import torch

from mfutil import HELD_OFF_GPU, assert_bitwise, eps, rand_mf, case_seed

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.fixture(scope="module")
def mfla():
    import mfla as M

    assert torch.cuda.is_available()
    return M


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_device_networks_equal_oracle(mfla, oracle, N):
    rng = np.random.default_rng(900 + N)
    a, b, c = (rand_mf(rng, N, 5000, spread=30) for _ in range(3))
    b[:, :300] = -a[:, :300]
    pos = a * np.sign(a[0])
    for op in ("add", "sub", "mul", "div", "mac", "msub", "inv"):
        got = host(mfla.arith(op, dev(a), dev(b), dev(c)))
        assert_bitwise(got, oracle.arith(op, a, b, c), f"x{N} {op}")
    assert_bitwise(host(mfla.arith("sqrt", dev(pos))), oracle.arith("sqrt", pos), f"x{N} sqrt")
    if N == 3:
        assert_bitwise(host(mfla.arith("mulacc_x3", dev(a), dev(b), dev(c))), oracle.arith("mulacc_x3", a, b, c), "fused")


def test_device_proof_vectors(mfla):
    import json
    import os

    with open(os.path.join(os.path.dirname(__file__), "golden", "mulacc_x3_proof_vectors.json")) as f:
        pv = json.load(f)
    f64 = lambda k: np.stack([np.array(v[k], dtype=np.uint64).view(np.float64) for v in pv], axis=1)  # noqa: E731
    acc, x, y, out = f64("acc"), f64("x"), f64("y"), f64("output")
    assert_bitwise(host(mfla.arith("mac", dev(x), dev(y), dev(acc))), out, "standard x3 mac")
    assert_bitwise(host(mfla.arith("mulacc_x3", dev(x), dev(y), dev(acc))), out, "mulacc_x3")


GEMM_SHAPES = [(1, 1, 1), (5, 3, 7), (33, 17, 31), (64, 64, 64), (70, 45, 129), (130, 9, 66), (4, 0, 5), (0, 3, 4)]


@pytest.mark.parametrize("N", [1, 2, 3, 4])
@pytest.mark.parametrize("shape", GEMM_SHAPES)
def test_gemm_matches_oracle(mfla, oracle, N, shape):
    m, k, n = shape
    rng = np.random.default_rng(m * 1000 + k * 10 + n + N)
    A, B, C = rand_mf(rng, N, m, k), rand_mf(rng, N, k, n), rand_mf(rng, N, m, n)
    alpha = rand_mf(rng, N, 1)[:, 0]
    beta = rand_mf(rng, N, 1)[:, 0]
    for a_, b_ in ((1.0, 0.0), (alpha, beta), (-1.0, 1.0)):
        strategies = ["auto", "direct", "fused"] if N == 3 else ["auto"]
        for strat in strategies:
            C0 = C.copy()
            if np.all(np.asarray(b_) == 0):
                C0[:] = np.nan  # beta == 0: the destination is never read
            Cd = dev(C0)
            mfla.gemm_(Cd, dev(A), dev(B), a_, b_, config=mfla.KernelConfig(gemm_strategy=strat))
            ref = oracle.gemm(C0.copy(), A, B, a_, b_, fused=(strat != "direct"))
            assert_bitwise(host(Cd), ref, f"x{N} {shape} {strat}")


@pytest.mark.parametrize("N", [1, 2, pytest.param(3, marks=HELD_OFF_GPU), 4])
def test_gemm_tile_shapes_bitwise_identical(mfla, oracle, N):
    """Every tile shape (16x16, 32x32, 32x64, 64x64) gives MFLA's bits. 131
    rows: for x3 the fused route runs mulacc_x3 on rows 0..127 and the
    standard network on the 3 remainder rows in a second launch."""
    rng = np.random.default_rng(5 + N)
    A, B = rand_mf(rng, N, 131, 40), rand_mf(rng, N, 40, 75)
    for strat in (("auto", "direct") if N == 3 else ("auto",)):
        ref = oracle.gemm(np.zeros((N, 75, 131)), A, B, 1.0, 0.0, fused=(strat == "auto"))
        for t in (0, 1, 2, 3):
            C = torch.zeros((N, 75, 131), dtype=torch.float64, device="cuda")
            mfla.gemm_(C, dev(A), dev(B), config=mfla.KernelConfig(gemm_tile=t, gemm_strategy=strat))
            assert_bitwise(host(C), ref, f"tile {t} x{N} {strat}")


def test_gemm_dimension_and_alias_errors(mfla):
    A = torch.zeros((2, 3, 4), dtype=torch.float64, device="cuda")
    B = torch.zeros((2, 5, 3), dtype=torch.float64, device="cuda")
    with pytest.raises(mfla.DimensionMismatch):
        mfla.gemm_(torch.zeros((2, 5, 5), dtype=torch.float64, device="cuda"), A, B)
    with pytest.raises(mfla.DimensionMismatch):
        mfla.gemm_(torch.zeros((2, 5, 4), dtype=torch.float64, device="cuda"), A, A)
    S = torch.zeros((2, 4, 4), dtype=torch.float64, device="cuda")
    with pytest.raises(mfla.ArgumentError):
        mfla.gemm_(S, S, S)
    with pytest.raises(mfla.ArgumentError):
        mfla.gemm_(torch.zeros((2, 5, 4), dtype=torch.float64, device="cuda"), A, B, config=mfla.KernelConfig(gemm_strategy="fused"))


@pytest.mark.parametrize("N", [1, 2, 3, 4])
@pytest.mark.parametrize("n,k", [(1, 1), (23, 5), (70, 33), (97, 16)])
def test_syrk_gemmt_lower_only(mfla, oracle, N, n, k):
    rng = np.random.default_rng(77 + n + k + N)
    P = rand_mf(rng, N, k, n)
    Lm, R = rand_mf(rng, N, n, k), rand_mf(rng, N, n, k)
    C = rand_mf(rng, N, n, n)
    iu = np.triu_indices(n, 1)
    C[:, iu[1], iu[0]] = np.nan           # upper triangle (row < col) must be preserved
    for a_, b_ in ((1.0, 0.0), (-1.0, 1.0), (rand_mf(rng, N, 1)[:, 0], rand_mf(rng, N, 1)[:, 0])):
        Cd = dev(C)
        mfla.syrk_(Cd, dev(P), a_, b_)
        got = host(Cd)
        ref = oracle.syrk(C.copy(), P, a_, b_)
        assert_bitwise(got, ref, f"syrk x{N}")
        assert np.all(np.isnan(got[:, iu[1], iu[0]]))
        Cd = dev(C)
        mfla.gemmt_(Cd, dev(Lm), dev(R), a_, b_)
        assert_bitwise(host(Cd), oracle.gemmt(C.copy(), Lm, R, a_, b_), f"gemmt x{N}")


def _triangular(rng, N, n, uplo, unit):
    A = rand_mf(rng, N, n, n)
    r, c = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")   # A[:, col, row]
    inactive = (c < r) if uplo == "lower" else (c > r)              # planar index [l, j, i]
    A[:, inactive] = 0.0
    idx = np.arange(n)
    A[0, idx, idx] = np.abs(A[0, idx, idx]) + 3.0          # well conditioned diagonal
    A[1:, idx, idx] = 0.0
    return A


@pytest.mark.parametrize("N", [2, 3, 4])
@pytest.mark.parametrize("side", ["left", "right"])
@pytest.mark.parametrize("uplo", ["lower", "upper"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["nonunit", "unit"])
def test_trsm_all_variants_exact_order(mfla, oracle, N, side, uplo, trans, diag):
    rng = np.random.default_rng(case_seed(N, side, uplo, trans, diag))
    n, other = 70, 9
    A = _triangular(rng, N, n, uplo, diag == "unit")
    B = rand_mf(rng, N, n, other) if side == "left" else rand_mf(rng, N, other, n)
    alpha = 1.0 if diag == "unit" else rand_mf(rng, N, 1)[:, 0]
    Bd = dev(B)
    mfla.trsm_(Bd, dev(A), alpha, side=side, uplo=uplo, trans=trans, diag=diag, config=mfla.KernelConfig(trsm_order="exact"))
    ref = oracle.trsm(B.copy(), A, alpha, side, uplo, trans, diag)
    assert_bitwise(host(Bd), ref, f"trsm x{N} {side} {uplo} {trans} {diag}")


@pytest.mark.parametrize("N,other", [(2, 7), (4, 7), (3, 7), (4, 70), (2, 70)])
@pytest.mark.parametrize("side,uplo,trans", [("left", "lower", "N"), ("left", "upper", "T"), ("right", "upper", "N"),
                                             ("right", "lower", "T")])
def test_trsm_forward_lookahead_bitwise(mfla, oracle, N, other, side, uplo, trans):
    """Forward orders at n = 330 (six 64-row blocks): the diagonal solves
    overlap the previous block's trailing update on a second stream (with 7
    right-hand sides and two or four limbs the next block's rows are updated
    inside each diagonal kernel; otherwise each diagonal kernel waits on a
    device counter for the update that last reaches its rows); every element
    keeps MFLA's ascending order, so the result is bitwise."""
    rng = np.random.default_rng(case_seed(N, other, side, uplo, trans))
    n = 330
    A = _triangular(rng, N, n, uplo, False)
    B = rand_mf(rng, N, n, other) if side == "left" else rand_mf(rng, N, other, n)
    Bd = dev(B)
    mfla.trsm_(Bd, dev(A), side=side, uplo=uplo, trans=trans)
    ref = oracle.trsm(B.copy(), A, 1.0, side, uplo, trans, "nonunit")
    assert_bitwise(host(Bd), ref, f"trsm look-ahead x{N} {side} {uplo} {trans}")


def test_trsm_lookahead_on_a_handle_made_after_many_streams(mfla, oracle):
    """Solves whose diagonal kernels wait on a device counter advanced on
    another stream, on a handle made once the process holds more streams than
    the runtime's hardware queues and issued on those streams (so the two
    streams may share a queue with each other or the caller's): each counter
    is advanced by work enqueued before the kernel that waits for it, so a
    shared queue runs that work first, and the results are those of the
    process's first handle (forward: bitwise MFLA's)."""
    import ctypes

    from mfla import _lib, api

    devno = torch.cuda.current_device()
    streams = [torch.cuda.Stream() for _ in range(6)]
    cases = [(2, 7, "N"), (3, 70, "N"), (4, 7, "T"), (2, 70, "T"), (4, 70, "N")]
    rng = np.random.default_rng(case_seed("many streams"))
    inputs, first = [], []
    for N, other, trans in cases:
        A = _triangular(rng, N, 330, "lower", False)
        B = rand_mf(rng, N, 330, other)
        Bd = dev(B)
        mfla.trsm_(Bd, dev(A), side="left", uplo="lower", trans=trans)
        inputs.append((A, B))
        first.append(host(Bd))
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.mfla_handle_create(ctypes.byref(h)) == 0
    old = api._handles.get(devno)
    try:
        api._handles[devno] = h
        for i, ((N, other, trans), (A, B)) in enumerate(zip(cases, inputs)):
            Ad, Bd = dev(A), dev(B)
            q = streams[i % len(streams)]
            q.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(q):
                mfla.trsm_(Bd, Ad, side="left", uplo="lower", trans=trans)
            torch.cuda.synchronize()
            assert_bitwise(host(Bd), first[i], f"trsm x{N} {other} RHS {trans} on a new handle")
            if trans == "N":
                ref = oracle.trsm(B.copy(), A, 1.0, "left", "lower", "N", "nonunit")
                assert_bitwise(host(Bd), ref, f"trsm x{N} {other} RHS forward")
    finally:
        api._handles[devno] = old
        assert lib.mfla_handle_destroy(h) == 0


@pytest.mark.parametrize("N,nrhs", [(2, 5), (4, 5), (3, 70)])
@pytest.mark.parametrize("uplo,trans", [("upper", "N"), ("lower", "T")])
def test_trsm_blocked_backward_within_tolerance(mfla, oracle, N, nrhs, uplo, trans):
    """Blocked backward order at n = 200 (four blocks, look-ahead: each
    diagonal kernel waits on a device counter for the side stream's update
    of its rows)."""
    rng = np.random.default_rng(31 + N + nrhs)
    n = 200
    A = _triangular(rng, N, n, uplo, False)
    B = rand_mf(rng, N, n, nrhs)
    Bd = dev(B)
    mfla.trsm_(Bd, dev(A), side="left", uplo=uplo, trans=trans, config=mfla.KernelConfig(trsm_order="blocked"))
    X = host(Bd)
    ref = oracle.trsm(B.copy(), A, 1.0, "left", uplo, trans, "nonunit")
    # blocked order rounds differently: check the exact residual B - op(A) X
    # with the long accumulator, and the exact difference from MFLA's order
    # against both solves' forward error bound c n u_N (|op(A)^-1| |op(A)| |x|)_i
    rows = np.repeat(np.arange(n), nrhs)
    cols = np.tile(np.arange(nrhs), n)
    At = np.ascontiguousarray(np.transpose(A, (0, 2, 1))) if trans == "T" else A
    val, scale = oracle.exact_axpb(At, X, B, sab=-1, sc=1, rows=rows, cols=cols)
    assert np.max(np.abs(val) / scale) <= 64 * n * eps(N)
    T = At[0].T                                                  # op(A), binary64 hi limbs, (row, col)
    skeel = np.abs(np.linalg.inv(T)) @ (np.abs(T) @ np.abs(ref[0].T))
    worst = 0.0
    for j in range(nrhs):
        for i in range(n):
            d = abs(oracle.exact_sum_diff(X[:, j, i], ref[:, j, i]))
            worst = max(worst, d / (skeel[i, j] * n * eps(N)))
    assert worst <= 2 * 64, worst


@pytest.mark.parametrize("N", [2, 3, 4])
def test_trsv_equals_one_column_trsm(mfla, oracle, N):
    # runtests.jl:683-690
    rng = np.random.default_rng(60 + N)
    n = 90
    for uplo in ("lower", "upper"):
        for trans in ("N", "T"):
            A = _triangular(rng, N, n, uplo, False)
            b = rand_mf(rng, N, n)
            x = dev(b)
            mfla.trsv_(x, dev(A), uplo=uplo, trans=trans)
            Bm = dev(b.reshape(N, 1, n))
            mfla.trsm_(Bm, dev(A), uplo=uplo, trans=trans)
            assert_bitwise(host(x), host(Bm).reshape(N, n), "trsv vs trsm")
            assert_bitwise(host(x), oracle.trsv(b.copy(), A, 1.0, uplo, trans), "trsv vs oracle")


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_gemv_symv(mfla, oracle, N):
    rng = np.random.default_rng(11 + N)
    m, n = 77, 45
    A = rand_mf(rng, N, m, n)
    for trans in ("N", "T"):
        x = rand_mf(rng, N, n if trans == "N" else m)
        y = rand_mf(rng, N, m if trans == "N" else n)
        for a_, b_ in ((1.0, 0.0), (-1.0, 1.0)):
            y0 = y.copy()
            if b_ == 0.0:
                y0[:] = np.nan
            yd = dev(y0)
            mfla.gemv_(yd, dev(A), dev(x), a_, b_, trans=trans)
            assert_bitwise(host(yd), oracle.gemv(y0.copy(), A, x, a_, b_, trans), f"gemv {trans}")
    S = rand_mf(rng, N, 60, 60)
    x = rand_mf(rng, N, 60)
    for uplo in ("lower", "upper"):
        S0 = S.copy()
        il = np.tril_indices(60, -1) if uplo == "upper" else np.triu_indices(60, 1)
        S0[:, il[1], il[0]] = np.nan       # inactive triangle is never read
        y = rand_mf(rng, N, 60)
        yd = dev(y)
        mfla.symv_(yd, dev(S0), dev(x), -1.0, 1.0, uplo=uplo)
        assert_bitwise(host(yd), oracle.symv(y.copy(), S0, x, -1.0, 1.0, uplo), f"symv {uplo}")


def test_gemm_x4_n1024_sampled_columns(mfla, oracle):
    # Full-size property at n = 1024: sampled output columns equal the oracle
    # bitwise (per-element results do not depend on other columns).
    rng = np.random.default_rng(1024)
    n = 1024
    A, B = rand_mf(rng, 4, n, n), rand_mf(rng, 4, n, n)
    Cd = torch.empty((4, n, n), dtype=torch.float64, device="cuda")
    mfla.gemm_(Cd, dev(A), dev(B))
    C = host(Cd)
    cols = [0, 1, 511, 1023]
    Bs = np.ascontiguousarray(B[:, cols, :])
    ref = oracle.gemm(np.zeros((4, len(cols), n)), A, Bs)
    assert_bitwise(C[:, cols, :], ref, "n=1024 sampled columns")
    val, scale = oracle.exact_axpb(A, Bs, ref, sab=1, sc=-1, rows=np.arange(0, n, 97).repeat(len(cols)),
                                   cols=np.tile(np.arange(len(cols)), len(range(0, n, 97))))
    assert np.max(np.abs(val) / scale) <= 8 * n * eps(4)
