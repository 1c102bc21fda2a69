"""mfla_sytrf_noalloc, the ldlt! entry point that allocates nothing, against
mfla_sytrf through the C ABI: factor, dsub, pivots, blocks, status and
original maximum bit for bit, on two calls in a row -- the second on a
smaller matrix, so it runs on the static device scratch the first call left
behind -- on the blocked route (n above the 512 crossover, the weighted
panel in the static scratch) and the unblocked one."""
import numpy as np
import pytest
import torch

from mfutil import assert_bitwise, rand_mf

pytestmark = pytest.mark.gpu


def symmetric_indefinite(rng, N, n):
    """A = G + G^T with G ~ rand_mf: indefinite, with 2x2 pivots and
    interchanges on the Bunch-Kaufman path."""
    G = rand_mf(rng, N, n, n)
    return G + np.transpose(G, (0, 2, 1))


def factor_both(M, A):
    ref = torch.from_numpy(A).cuda()
    got = torch.from_numpy(A).cuda()
    F = M.ldlt_(ref, check=False)
    G = M.ldlt_noalloc_(got, check=False)
    torch.cuda.synchronize()
    return F, G, ref, got


@pytest.mark.parametrize("N", [2, 4])
@pytest.mark.parametrize("sizes", [(700, 530), (300, 170)])
def test_noalloc_ldlt_equals_sytrf_on_two_calls(N, sizes):
    import mfla as M

    rng = np.random.default_rng(1000 * N + sizes[0])
    for n in sizes:
        print(f"noalloc ldlt x{N} n={n}", flush=True)
        F, G, ref, got = factor_both(M, symmetric_indefinite(rng, N, n))
        assert F.info == G.info == 0, (F.info, G.info)
        assert np.array_equal(F.pivots.cpu().numpy(), G.pivots.cpu().numpy()), n
        assert np.array_equal(F.blocks.cpu().numpy(), G.blocks.cpu().numpy()), n
        assert (F.blocks.cpu().numpy() == 2).any()
        assert_bitwise(G.dsub.cpu().numpy(), F.dsub.cpu().numpy(), f"dsub n={n}")
        assert_bitwise(got.cpu().numpy(), ref.cpu().numpy(), f"factor n={n}")
        assert_bitwise(G.original_maximum, F.original_maximum, f"original maximum n={n}")


def test_noalloc_ldlt_failure_state_equals_sytrf():
    """A singular matrix (a zero row and column): the same failing step and
    the same partially factored matrix."""
    import mfla as M

    rng = np.random.default_rng(77)
    for n, zero in ((600, 333), (200, 50)):
        A = symmetric_indefinite(rng, 4, n)
        A[:, zero, :] = 0.0
        A[:, :, zero] = 0.0
        F, G, ref, got = factor_both(M, A)
        assert F.info == G.info and F.info > 0, (F.info, G.info)
        assert_bitwise(got.cpu().numpy(), ref.cpu().numpy(), f"failed factor n={n}")
        assert_bitwise(G.dsub.cpu().numpy(), F.dsub.cpu().numpy(), f"dsub n={n}")
