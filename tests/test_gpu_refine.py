"""FP64-factor + Float64x4 refinement (a build extension: MFLA has no FP64 ->
x4 refinement loop). Pinned by the exact long accumulator: the refined
solution's true normwise backward error (residual evaluated exactly, not in
Float64x4) must be <= 1e-60, like the plain Float64x4 factor + solve."""
import numpy as np
import pytest
import torch

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
    # the convergence is geometric at ~kappa * u_64 per step
    assert hist[1] < 1e-6 * hist[0]
    Ah, Xh = A.cpu().numpy(), X.cpu().numpy()
    be = exact_backward_error(oracle, Ah, Xh, Bh)
    assert be.max() <= 1e-60, be
    Xp = R.plain_solve(A, B, kind).cpu().numpy()
    bp = exact_backward_error(oracle, Ah, Xp, Bh)
    assert bp.max() <= 1e-60, bp
