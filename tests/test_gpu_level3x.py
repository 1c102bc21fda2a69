"""trmm!, mfdot and syrk_packed! on the GPU: bitwise equal to the oracle
(MFLA's per-element order), inactive triangles never read, MFLA's errors."""
import itertools

import numpy as np
import pytest
import torch

from mfutil import assert_bitwise, rand_mf, case_seed

pytestmark = pytest.mark.gpu


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("side,uplo,trans,diag", list(itertools.product(("left", "right"), ("lower", "upper"), ("N", "T"),
                                                                         ("nonunit", "unit"))))
@pytest.mark.parametrize("N", [2, 3, 4])
def test_trmm_bitwise(oracle, side, uplo, trans, diag, N):
    from mfla import api as M

    rng = np.random.default_rng(case_seed(side, uplo, trans, diag, N))
    n, r = 45, 37
    A = rand_mf(rng, N, n, n)
    inactive = np.triu_indices(n, 1) if uplo == "lower" else np.tril_indices(n, -1)
    Ag = A.copy()
    Ag[:, inactive[1], inactive[0]] = np.nan                      # planar [col, row]
    B = rand_mf(rng, N, n, r) if side == "left" else rand_mf(rng, N, r, n)
    alpha = rand_mf(rng, N, 1)[:, 0]
    ref = oracle.trmm(B.copy(), A, alpha, side, uplo, trans, diag)
    Bg = cu(B)
    M.trmm_(Bg, cu(Ag), alpha, side=side, uplo=uplo, trans=trans, diag=diag)
    assert_bitwise(Bg.cpu().numpy(), ref, f"trmm {side} {uplo} {trans} {diag} x{N}")


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_dot_bitwise(oracle, N):
    from mfla import api as M

    rng = np.random.default_rng(N)
    for rows, cols in ((1, 1), (1000, 1), (33, 70), (2500, 3)):
        x, y = rand_mf(rng, N, rows, cols), rand_mf(rng, N, rows, cols)
        assert_bitwise(M.mfdot(cu(x), cu(y)), oracle.dot(x, y), f"dot {rows}x{cols}")
    with pytest.raises(M.DimensionMismatch):
        M.mfdot(cu(rand_mf(rng, N, 3, 2)), cu(rand_mf(rng, N, 2, 3)))


@pytest.mark.parametrize("N", [2, 3, 4])
def test_syrk_packed_bitwise(oracle, N):
    from mfla import api as M

    rng = np.random.default_rng(60 + N)
    n, k = 70, 41
    P = rand_mf(rng, N, k, n)
    for rf, rl, alpha, beta in ((1, k, 1.0, 0.0), (5, 30, -0.75, 1.0)):
        packed = rand_mf(rng, N, n * (n + 1) // 2) if beta != 0.0 else np.full((N, n * (n + 1) // 2), np.nan)
        ref = oracle.syrk_packed(packed.copy() if beta != 0.0 else np.zeros_like(packed), P, alpha, beta, rf, rl)
        g = cu(packed)
        M.syrk_packed_(g, cu(P), alpha, beta, reduction_first=rf, reduction_last=rl)
        assert_bitwise(g.cpu().numpy(), ref, f"syrk_packed rf={rf} rl={rl}")
    with pytest.raises(M.BoundsError):
        M.syrk_packed_(g, cu(P), reduction_first=0, reduction_last=3)
    with pytest.raises(M.DimensionMismatch):
        M.syrk_packed_(g[:, :-1].contiguous(), cu(P))


def test_trmm_errors():
    from mfla import api as M

    A = torch.zeros((2, 4, 4), dtype=torch.float64, device="cuda")
    with pytest.raises(M.DimensionMismatch):
        M.trmm_(torch.zeros((2, 3, 5), dtype=torch.float64, device="cuda"), A)
    with pytest.raises(M.ArgumentError):
        M.trmm_(torch.zeros((2, 3, 4), dtype=torch.float64, device="cuda"), A, side="middle")
    with pytest.raises(M.ArgumentError):
        M.trmm_(A[:, :, :], A)                                     # output aliases the triangle


def test_capabilities_table():
    from mfla import api as M

    c = M.capabilities(4)
    assert c["provider"] == "mfla_hip" and c["supported"] and c["trmm"] and c["syrk_packed"] and c["dot"]
    assert c["mixed_residual_targets"] == {"x2": False, "x3": False, "x4": False}
    assert M.capabilities(2)["mixed_residual_target_types"] == ("Float64x3", "Float64x4")
    assert not M.capabilities(5)["supported"] and not M.capabilities(5)["gemm"]
