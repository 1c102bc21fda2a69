"""GPU parity at the full sizes of the benchmark configurations.

Per-element results do not depend on other output columns (ascending k per
element, no split-K), so sampled output columns of a full-size launch are
compared bitwise with the oracle run on just those columns; a row sample of
each is also bounded against the exact long accumulator.

* gemm! square n = 2048 and 4096 and tall-skinny 8192 x 512 x 8192 for
  Float64x2/x3/x4 (the x3 auto route runs mulacc_x3 in the 32 x 32 tile);
  the x3 fused 32 x 32 tile together with its m mod 4 remainder launch;
* syrk! n = 4096, k = 1024 (sampled lower columns; upper triangle untouched);
* trsm! lower 4096 x 4096 with 256 right-hand sides on the Cholesky factor of
  R R^T + n I (forward order: bitwise);
* the refined and the plain Float64x4 solve at n = 4096, kappa = 1e8:
  exact backward error <= 1e-60;
* a deterministic sample of the 10,000-system batch: factor, pivots, blocks
  and inertia bitwise against the unblocked oracle, solutions by the exact
  residual (and bitwise in MFLA's exact backward order).
"""
import numpy as np
import pytest
import torch

from mfutil import HELD_OFF_GPU, assert_bitwise, eps

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mfla():
    import mfla as M

    assert torch.cuda.is_available()
    return M


def rand_dev(N, rows, cols, seed):
    """Seeded Float64xN on the device: hi ~ N(0, 1), limb l = limb(l-1) *
    2^-53 * U(-1, 1), renormalised by two fast_two_sum sweeps."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.empty((N, cols, rows), dtype=torch.float64, device="cuda")
    x[0] = torch.randn((cols, rows), generator=g, dtype=torch.float64, device="cuda")
    for l in range(1, N):
        x[l] = x[l - 1] * 2.0**-53 * (torch.rand((cols, rows), generator=g, dtype=torch.float64, device="cuda") * 2 - 1)
    for _ in range(2):
        for l in range(N - 1):
            s = x[l] + x[l + 1]
            e = x[l + 1] - (s - x[l])
            x[l], x[l + 1] = s, e
    return x


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def sample_cols(n):
    return sorted({0, 1, n // 2 + 3, n - 1})


@pytest.mark.parametrize("N", [2, 3, 4])
@pytest.mark.parametrize("m,k,n", [(2048, 2048, 2048), (4096, 4096, 4096), (8192, 512, 8192)])
def test_gemm_config_sizes_sampled_columns(mfla, oracle, N, m, k, n):
    print(f"gemm x{N} {m}x{k}x{n}", flush=True)
    A, B = rand_dev(N, m, k, 11), rand_dev(N, k, n, 12)
    C = torch.empty((N, n, m), dtype=torch.float64, device="cuda")
    mfla.gemm_(C, A, B)
    cols = sample_cols(n)
    got = host(C[:, cols, :])
    del C
    Ah, Bs = host(A), host(B[:, cols, :].contiguous())
    ref = oracle.gemm(np.zeros((N, len(cols), m)), Ah, Bs)          # auto route: mulacc_x3 rows for x3
    assert_bitwise(got, ref, f"gemm x{N} {m}x{k}x{n} sampled columns")
    rows = np.arange(0, m, 61)
    val, scale = oracle.exact_axpb(Ah, Bs, ref, sab=1, sc=-1, rows=np.repeat(rows, len(cols)),
                                   cols=np.tile(np.arange(len(cols)), len(rows)))
    assert np.max(np.abs(val) / scale) <= 8 * k * eps(N)


@pytest.mark.parametrize("strategy", ["auto", "direct"])
def test_gemm_x3_32x32_tile_with_remainder_rows(mfla, oracle, strategy):
    """731 x 70 x 720: the automatic choice is the 32 x 32 tile; on the fused
    route rows 0..727 run mulacc_x3 and the last 3 rows the standard network
    in a second launch (MFLA's four-row vector blocks, gemm.jl:419-451)."""
    N, m, k, n = 3, 731, 70, 720
    A, B = rand_dev(N, m, k, 21), rand_dev(N, k, n, 22)
    for alpha, beta in ((1.0, 0.0), (-1.0, 1.0)):
        C0 = rand_dev(N, m, n, 23)
        C = C0.clone()
        mfla.gemm_(C, A, B, alpha, beta, config=mfla.KernelConfig(gemm_strategy=strategy))
        ref = oracle.gemm(host(C0), host(A), host(B), alpha, beta, fused=(strategy == "auto"))
        assert_bitwise(host(C), ref, f"x3 {strategy} alpha={alpha} beta={beta}")


@pytest.mark.parametrize("N", [2, 3, 4])
def test_syrk_config_size_sampled_columns(mfla, oracle, N):
    print(f"syrk x{N} n=4096 k=1024", flush=True)
    n, k = 4096, 1024
    P = rand_dev(N, k, n, 31)                                         # k x n
    C = torch.full((N, n, n), float("nan"), dtype=torch.float64, device="cuda")
    mfla.syrk_(C, P)                                                  # beta = 0: C never read
    Ph = host(P)
    Pt = np.ascontiguousarray(np.transpose(Ph, (0, 2, 1)))           # P^T, n x k
    for c in sample_cols(n):
        col = host(C[:, c, :])
        assert np.all(np.isnan(col[:, :c])), f"upper triangle of column {c} was written"
        # lower column c: acc = sum_k P[k, r] P[k, c], r >= c -- the standard network, as gemm's direct route
        ref = oracle.gemm(np.zeros((N, 1, n - c)), np.ascontiguousarray(Pt[:, :, c:]), np.ascontiguousarray(Ph[:, c:c + 1, :]),
                          fused=False)
        assert_bitwise(col[:, c:], ref[:, 0, :], f"syrk x{N} column {c}")


@pytest.mark.parametrize("N", [2, 3, 4])
def test_trsm_config_size_sampled_rhs(mfla, oracle, N):
    print(f"trsm x{N} 4096 x 256", flush=True)
    n, nrhs = 4096, 256
    R = rand_dev(N, n, n, 41)
    S = torch.zeros_like(R)
    mfla.gemm_(S, R, R.transpose(1, 2).contiguous())
    del R
    S[0].diagonal().add_(float(n))
    F = mfla.cholesky_(S)
    assert F.info == 0
    B0 = rand_dev(N, n, nrhs, 42)
    X = B0.clone()
    mfla.trsm_(X, F.factors)                                          # left, lower, N: forward order
    cols = sample_cols(nrhs)
    got = host(X[:, cols, :])
    L = host(F.factors)
    Bs = host(B0[:, cols, :].contiguous())
    ref = oracle.trsm(Bs.copy(), L, 1.0, "left", "lower", "N", "nonunit")
    assert_bitwise(got, ref, f"trsm x{N} sampled right-hand sides")
    Lt = L.copy()
    iu = np.triu_indices(n, 1)
    Lt[:, iu[1], iu[0]] = 0.0                                         # only the lower triangle is the factor
    rows = np.arange(0, n, 37)
    val, scale = oracle.exact_axpb(Lt, ref, Bs, sab=-1, sc=1, rows=np.repeat(rows, len(cols)),
                                   cols=np.tile(np.arange(len(cols)), len(rows)))
    assert np.max(np.abs(val) / scale) <= 8 * n * eps(N)


def exact_backward_error(oracle, A, X, B, cols):
    """max_i |b - A x|_i (exact) / (|A|_inf |x|_inf + |b|_inf) per sampled column."""
    N, n, _ = A.shape
    Xs = np.ascontiguousarray(X[:, cols, :])
    Bs = np.ascontiguousarray(B[:, cols, :])
    rows = np.arange(n)
    val, _ = oracle.exact_axpb(A, Xs, Bs, sab=-1, sc=1, rows=np.repeat(rows, len(cols)),
                               cols=np.tile(np.arange(len(cols)), n))
    rn = np.abs(val).reshape(n, len(cols)).max(axis=0)
    an = np.abs(A.sum(0)).sum(axis=0).max()
    xn = np.abs(Xs.sum(0)).max(axis=1)
    bn = np.abs(Bs.sum(0)).max(axis=1)
    return rn / (an * xn + bn)


@pytest.mark.parametrize("kind", ["cholesky", "lu"])
def test_refined_and_plain_solve_n4096_exact_backward_error(mfla, oracle, kind):
    from mfla import refine as R

    print(f"refine {kind} n=4096", flush=True)
    n, nrhs = 4096, 64
    A = R.spd_kappa(n, 1e8) if kind == "cholesky" else R.general_kappa(n, 1e8)
    rng = np.random.default_rng(7)
    Bh = np.zeros((4, nrhs, n))
    Bh[0] = rng.standard_normal((nrhs, n))
    for l in range(1, 4):
        Bh[l] = Bh[l - 1] * 2.0**-53 * rng.uniform(-1, 1, (nrhs, n))
    B = torch.from_numpy(Bh).cuda()
    X, hist = R.refine_solve(A, B, kind)
    assert hist[-1] <= 2.0**-216, hist
    Ah = host(A)
    cols = [0, 21, 42, 63]
    be = exact_backward_error(oracle, Ah, host(X), Bh, cols)
    assert be.max() <= 1e-60, be
    Xp = R.plain_solve(A, B, kind)
    bp = exact_backward_error(oracle, Ah, host(Xp), Bh, cols)
    assert bp.max() <= 1e-60, bp


SAMPLE = list(range(0, 10000, 100)) + list(range(1, 10000, 100))


def config3_batch():
    from mfla.batched import synthetic_systems

    return 3, synthetic_systems(10000, 16, 256, 3, 4, seed=3)


def test_batched_config3_sample(mfla, oracle):
    """Every 100th system of the 10,000-system x3 batch (n ~ U{16..256},
    SPD / KKT alternating, 4 right-hand sides), solved as the benchmark
    solves it (systolic backward order): factor, pivots, blocks, D and
    inertia bitwise against the oracle, X by the exact residual."""
    from backends import GpuBackend, OracleBackend
    from mfla.batched import Batch
    from test_gpu_batched import check_solution_residual

    ob, gb = OracleBackend(oracle), GpuBackend()
    N, systems = config3_batch()
    b = Batch(systems, N)
    b.solve()
    res = b.results()
    for idx in SAMPLE:
        kind, A, B = systems[idx]
        r = res[idx]
        F = A.copy()
        if kind == 0:
            info = oracle.potrf(F, 16)
            assert r["info"] == info == 0, idx
            assert_bitwise(r["A"], F, f"system {idx} cholesky")
        else:
            info, dsub, piv, blk, om = oracle.sytrf(F, "unblocked")
            assert r["info"] == info == 0, idx
            assert np.array_equal(r["pivots"], piv) and np.array_equal(r["blocks"], blk), idx
            assert_bitwise(r["dsub"], dsub, f"system {idx} dsub")
            assert_bitwise(r["A"], F, f"system {idx} ldlt")
            # Sylvester: [H A'; A -1e-8 I] with H SPD (m x m) has inertia (m, p, 0)
            n = A.shape[1]
            m = int(round(2 * n / 3))
            assert ob.inertia(F, dsub, blk) == (m, n - m, 0), idx
            assert gb.inertia(r["A"], r["dsub"], r["blocks"]) == (m, n - m, 0), idx
        check_solution_residual(oracle, N, kind, A, B, r["X"])


@HELD_OFF_GPU
def test_batched_config3_sample_exact_backward(mfla, oracle):
    """The same sample with MFLA's exact backward order: X bitwise."""
    from mfla.batched import Batch

    N, systems = config3_batch()
    b = Batch(systems, N)
    b.solve(exact_backward=True)
    res = b.results()
    for idx in SAMPLE:
        kind, A, B = systems[idx]
        F = A.copy()
        if kind == 0:
            assert oracle.potrf(F, 16) == 0
            assert_bitwise(res[idx]["X"], oracle.potrs(F, B.copy()), f"system {idx} potrs")
        else:
            info, dsub, piv, blk, _ = oracle.sytrf(F, "unblocked")
            assert info == 0
            assert_bitwise(res[idx]["X"], oracle.sytrs(F, dsub, piv, blk, B.copy()), f"system {idx} sytrs")
