"""Batched persistent factor + solve: every system bitwise equal to the oracle
(factor, pivots, blocks, D, status), solutions bitwise in MFLA's exact order,
systolic-order solutions within c*n*u of the exact residual, and a captured
HIP graph replayed several times (with eager work in between) reproducing the
eager result."""
import numpy as np
import pytest
import torch

from mfutil import HELD_OFF_GPU, assert_bitwise, eps, rand_mf

pytestmark = pytest.mark.gpu


def make_batch(oracle, rng, N, count, nmin, nmax, nrhs=4):
    systems = []
    for i in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        if i % 2 == 0:
            R = rand_mf(rng, N, n, n)
            S = np.zeros((N, n, n))
            oracle.gemm(S, R, np.ascontiguousarray(np.transpose(R, (0, 2, 1))), fused=False)
            S[0, np.arange(n), np.arange(n)] += n
            systems.append((0, S, rand_mf(rng, N, n, nrhs)))
        else:
            m = max(2, (2 * n) // 3)
            p = n - m
            H = rand_mf(rng, N, m, m)
            K = np.zeros((N, n, n))
            K[:, :m, :m] = H + np.transpose(H, (0, 2, 1))
            K[0, np.arange(m), np.arange(m)] += 2.0 * m
            A = rand_mf(rng, N, p, m)
            K[:, :m, m:] = A
            K[:, m:, :m] = np.transpose(A, (0, 2, 1))
            K[0, np.arange(m, n), np.arange(m, n)] = -1e-8
            systems.append((1, K, rand_mf(rng, N, n, nrhs)))
    return systems


def weak_kkt(rng, N, n, nrhs=4):
    """KKT [H A'; A -1e-8 I] whose H is small against A (H = 1e-3 (G + G'),
    G ~ N(0, 1)): Bunch-Kaufman pivots with 2x2 blocks and interchanges that
    reach the constraint rows, i.e. the last third of the matrix."""
    m = (2 * n) // 3
    p = n - m
    H = rand_mf(rng, N, m, m) * 1e-3
    K = np.zeros((N, n, n))
    K[:, :m, :m] = H + np.transpose(H, (0, 2, 1))
    A = rand_mf(rng, N, p, m)
    K[:, :m, m:] = A
    K[:, m:, :m] = np.transpose(A, (0, 2, 1))
    K[0, np.arange(m, n), np.arange(m, n)] = -1e-8
    return (1, K, rand_mf(rng, N, n, nrhs))


def check_solution_residual(oracle, N, kind, A, B, X):
    """X of the systolic backward order (rounds differently from MFLA's exact
    order): the residual B - A X summed exactly, within c*n*u_N."""
    n, nrhs = A.shape[1], B.shape[1]
    Af = A.copy()
    if kind == 0:                       # symmetric from the lower triangle
        il = np.triu_indices(n, 1)
        Af[:, il[1], il[0]] = Af[:, il[0], il[1]]
    rows = np.repeat(np.arange(n), nrhs)
    cols = np.tile(np.arange(nrhs), n)
    val, scale = oracle.exact_axpb(Af, X, B, sab=-1, sc=1, rows=rows, cols=cols)
    assert np.max(np.abs(val) / scale) <= 4096 * n * eps(N), n


def check_against_oracle(oracle, systems, res, exact_x=True, N=None):
    """Factor, status, pivots, blocks, D and original maximum bitwise; X
    bitwise in MFLA's exact order (exact_x) or by the exact residual."""
    for (kind, A, B), r in zip(systems, res):
        n = A.shape[1]
        if kind == 0:
            F = A.copy()
            info = oracle.potrf(F, 16)
            assert r["info"] == info == 0
            assert_bitwise(r["A"], F, f"batched cholesky n={n}")
            if exact_x:
                assert_bitwise(r["X"], oracle.potrs(F, B.copy()), f"batched potrs n={n}")
        else:
            F = A.copy()
            info, dsub, piv, blk, om = oracle.sytrf(F, "unblocked")
            assert r["info"] == info == 0
            assert np.array_equal(r["pivots"], piv) and np.array_equal(r["blocks"], blk), n
            assert_bitwise(r["dsub"], dsub, "dsub")
            assert_bitwise(r["A"], F, f"batched ldlt n={n}")
            assert_bitwise(r["original_maximum"], om, "original maximum")
            if exact_x:
                assert_bitwise(r["X"], oracle.sytrs(F, dsub, piv, blk, B.copy()), f"batched sytrs n={n}")
        if not exact_x:
            check_solution_residual(oracle, N, kind, A, B, r["X"])


@pytest.mark.parametrize("N", [3, 4])
def test_batched_bitwise_and_exact(oracle, N):
    from mfla.batched import Batch

    rng = np.random.default_rng(400 + N)
    systems = make_batch(oracle, rng, N, 24, 3, 90)
    b = Batch(systems, N)
    b.solve(exact_backward=True)
    check_against_oracle(oracle, systems, b.results())


def orders_91_to_256_systems(oracle):
    """The configuration's full order range: n in [91, 256] (n = 256
    included), so the workgroup's 256 threads all own rows; the weak-H KKT
    systems take interchanges with rows past 128, i.e. candidates found in
    waves 2-3 of the argmax."""
    N = 3
    rng = np.random.default_rng(4096)
    systems = make_batch(oracle, rng, N, 8, 91, 255)
    systems += make_batch(oracle, rng, N, 2, 256, 256)
    systems += [weak_kkt(rng, N, n) for n in (256, 227, 160)]
    return N, systems


def test_batched_orders_91_to_256_systolic(oracle):
    """Factor, status, pivots, blocks, D bitwise and inertia-relevant
    interchanges past row 128; X (systolic backward order, the default) by
    the exact residual."""
    from mfla.batched import Batch

    N, systems = orders_91_to_256_systems(oracle)
    b = Batch(systems, N)
    b.solve()
    res = b.results()
    check_against_oracle(oracle, systems, res, exact_x=False, N=N)
    for (kind, A, _), r in zip(systems[-3:], res[-3:]):
        n = A.shape[1]
        moved = np.abs(r["pivots"]) != np.arange(1, n + 1)
        assert moved.any() and (np.abs(r["pivots"][moved]) > 128).any(), (n, r["pivots"][:20])
        assert (r["blocks"] == 2).any()


@HELD_OFF_GPU
def test_batched_orders_91_to_256_exact_backward(oracle):
    """The same systems with MFLA's exact backward order: X bitwise."""
    from mfla.batched import Batch

    N, systems = orders_91_to_256_systems(oracle)
    b = Batch(systems, N)
    b.solve(exact_backward=True)
    check_against_oracle(oracle, systems, b.results())


def test_batched_systolic_order_and_graph(oracle):
    from mfla.batched import Batch, autotune

    N = 3
    rng = np.random.default_rng(77)
    systems = make_batch(oracle, rng, N, 40, 16, 128)
    b = Batch(systems, N)
    tune = autotune(b, shape_sets=((256,), (64, 128, 256)), scales=(0.25, 1), use_cache=False)
    assert tune["shapes"] in ((256,), (64, 128, 256)) and len(tune["trials"]) == 8
    b.restore()
    b.solve(tune["workgroups"])
    eager = b.results()
    for (kind, A, B), r in zip(systems, eager):
        assert r["info"] == 0
        check_solution_residual(oracle, N, kind, A, B, r["X"])
    g = b.graph(tune["workgroups"])
    for rep in range(3):
        junk = torch.randn(1000, device="cuda") * 2          # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        again = b.results()
        for r0, r1 in zip(eager, again):
            assert r0["info"] == r1["info"]
            assert_bitwise(r1["X"], r0["X"], f"graph replay {rep}")
            assert_bitwise(r1["A"], r0["A"], f"graph replay {rep}")
        del junk


@pytest.mark.parametrize("shapes", [(64, 256), (128, 256), (64, 128, 256)])
def test_batched_workgroup_shapes_bitwise_identical(oracle, shapes):
    """Size classes on 64 / 128 / 256-thread workgroups: every output of every
    system bitwise equal to the single 256-thread launch, and a sample equal
    to the oracle."""
    from mfla.batched import Batch

    N = 3
    rng = np.random.default_rng(9)
    systems = make_batch(oracle, rng, N, 30, 3, 200)
    one = Batch(systems, N)
    one.solve()
    ref = one.results()
    b = Batch(systems, N, shapes=shapes)
    assert len(b.classes) >= 2
    b.solve()
    got = b.results()
    for i, (r0, r1) in enumerate(zip(ref, got)):
        assert r0["info"] == r1["info"] == 0, i
        for k in ("A", "X", "original_maximum", "pivots", "blocks", "dsub"):
            if k in r0:
                assert np.array_equal(np.asarray(r0[k]).view(np.uint8), np.asarray(r1[k]).view(np.uint8)), (i, k)
    check_against_oracle(oracle, systems[:6], got[:6], exact_x=False, N=N)


def test_batched_rejects_oversize_and_reports_failures(oracle):
    from mfla.batched import Batch

    N = 2
    bad = np.zeros((N, 3, 3))
    bad[0] = np.array([[1.0, 2, 0], [2, 1, 0], [0, 0, 1]])      # not SPD: fails at 2
    sing = np.ones((N, 2, 2))
    sing[1:] = 0.0
    b = Batch([(0, bad, np.ones((N, 1, 3))), (1, sing, np.ones((N, 1, 2)))], N)
    b.solve()
    r = b.results()
    assert r[0]["info"] == 2 and r[1]["info"] == 2
