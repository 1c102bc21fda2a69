"""FP64-factor + Float64xN iterative refinement, and its benchmark.

MFLA provides the pieces -- residual!, refinement_correction! (exactly one
ldiv!), normwise_backward_error -- and leaves the loop to the caller
(src/residual.jl:294-300, docs/SDPX_PROVIDER_READINESS.md). The library's
native driver (mfla_refine_solve) runs that loop on the device:

    F64 = factor(hi limb of A), inv(F64) by recursive doubling
    X = 0, R = B
    repeat:  D = inv(F64) hi(R)    (two binary64 GEMMs on the matrix cores)
             X = X + D             (Float64xN accumulation)
             R = R - A*D           (summed with just enough limbs for |D|/|X|)
             until the backward error estimate <= tol or it stops improving

Convergence needs kappa(A) * u_64 < 1 (kappa <= 1e8 here); each step gains
~ -log10(kappa * u_64) digits until the Float64xN residual floor. The same
loop over MFLA's public pieces is refine_solve_pieces.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import api as M


def _eye(N, n):
    Q = torch.zeros((N, n, n), dtype=torch.float64, device="cuda")
    Q[0].diagonal().fill_(1.0)
    return Q


def householder_orthogonal(N, n, reflectors, rng):
    """Q = H_1 ... H_r, H = I - tau v v^T with v ~ N(0, 1) (FP64 values) and
    tau = 2 / (v^T v) in Float64xN, formed with gemm! in Float64xN."""
    Q = _eye(N, n)
    for _ in range(reflectors):
        v = torch.zeros((N, 1, n), dtype=torch.float64, device="cuda")
        v[0, 0] = torch.from_numpy(rng.standard_normal(n)).cuda()
        vt = v.reshape(N, n, 1).contiguous()                       # 1 x n
        vv = torch.zeros((N, 1, 1), dtype=torch.float64, device="cuda")
        M.gemm_(vv, vt, v)
        two = np.zeros((N, 1))
        two[0, 0] = 2.0
        tau = M.host_arith("div", two, vv.reshape(N, 1).cpu().numpy())[:, 0]
        w = torch.zeros((N, 1, n), dtype=torch.float64, device="cuda")
        M.gemm_(w, Q, v)
        M.gemm_(Q, w, vt, -tau, 1.0)
    return Q


def haar_orthogonal(N, n, rng):
    """A Haar-random orthogonal Q: the binary64 Householder QR of a seeded
    n x n N(0, 1) matrix (LAPACK dgeqrf / dorgqr through numpy: Q is the
    product of n - 1 non-trivial reflectors), each column signed so that R's
    diagonal is positive. Returned as Float64xN with zero lower limbs, i.e.
    the binary64 values exactly (orthogonal to ~1e-15)."""
    G = rng.standard_normal((n, n))
    Q, Rr = np.linalg.qr(G)
    Q *= np.where(np.diagonal(Rr) < 0, -1.0, 1.0)[None, :]
    out = torch.zeros((N, n, n), dtype=torch.float64, device="cuda")
    out[0] = torch.from_numpy(np.ascontiguousarray(Q.T)).cuda()     # planar [col, row]
    return out


# Orthogonal factors of the conditioned test matrices: "householder4" (the
# product of 4 random reflectors; the default) or "haar" (Haar-random).
ORTHOGONAL = "householder4"
_FACTORS = {"householder4": "4 random Householder reflectors (I - tau v v^T, v ~ N(0, 1))",
            "haar": "Haar-random, n - 1 reflectors (binary64 Householder QR of an n x n Gaussian)"}


def orthogonal(N, n, rng, kind=None):
    kind = kind or ORTHOGONAL
    if kind == "haar":
        return haar_orthogonal(N, n, rng)
    if kind == "householder4":
        return householder_orthogonal(N, n, 4, rng)
    raise ValueError("orthogonal factors must be 'householder4' or 'haar'")


def _scale_columns(Q, sigma):
    """Q * diag(sigma) elementwise in Float64xN (sigma: FP64 values)."""
    N, n, _ = Q.shape
    S = torch.zeros_like(Q)
    S[0] = torch.from_numpy(np.ascontiguousarray(np.repeat(sigma[:, None], n, axis=1))).cuda()
    return M.arith("mul", Q.reshape(N, -1), S.reshape(N, -1)).reshape(N, n, n)


def _symmetrize_lower(A):
    N, n, _ = A.shape
    idx = torch.arange(n, device="cuda")
    lower = (idx[None, :] >= idx[:, None])          # planar [col, row]: row >= col
    return torch.where(lower[None], A, A.transpose(1, 2)).contiguous()


def spd_kappa(n, kappa=1e8, seed=20260812, N=4, factors=None):
    """A = Q diag(sigma) Q^T, sigma log-spaced over [1, kappa], Q from
    orthogonal(factors), formed in Float64xN and made exactly symmetric."""
    rng = np.random.default_rng(seed)
    Q = orthogonal(N, n, rng, factors)
    sigma = 10.0 ** (np.log10(kappa) * np.arange(n) / max(n - 1, 1))
    W = _scale_columns(Q, sigma)
    A = torch.zeros_like(Q)
    M.gemm_(A, W, Q.transpose(1, 2).contiguous())
    return _symmetrize_lower(A)


def general_kappa(n, kappa=1e8, seed=20260813, N=4, factors=None):
    """A = U diag(sigma) V^T with independent U, V from orthogonal(factors)."""
    rng = np.random.default_rng(seed)
    U = orthogonal(N, n, rng, factors)
    V = orthogonal(N, n, rng, factors)
    sigma = 10.0 ** (np.log10(kappa) * np.arange(n) / max(n - 1, 1))
    W = _scale_columns(U, sigma)
    A = torch.zeros_like(U)
    M.gemm_(A, W, V.transpose(1, 2).contiguous())
    return A


def _max_be(be):
    return float(np.max(np.abs(be[0]))) if be.size else 0.0


class _Timer:
    """Accumulates device time per stage with CUDA events (optional)."""

    def __init__(self, on):
        self.on, self.t, self.ev = on, {}, []

    def mark(self, name):
        if self.on:
            e = torch.cuda.Event(enable_timing=True)
            e.record()
            self.ev.append((name, e))

    def done(self):
        if not self.on:
            return None
        torch.cuda.synchronize()
        for (_, a), (name, b) in zip(self.ev, self.ev[1:]):
            self.t[name] = self.t.get(name, 0.0) + a.elapsed_time(b)
        return {k: round(v, 3) for k, v in self.t.items()}


_REFINE_KIND = {"cholesky": 0, "lu": 1}
_STAGES = ("factor", "solve", "update", "residual", "check")
_ws_cache: dict = {}


def _refine_workspace(N, kind, n, nrhs, flag=0):
    import ctypes
    from . import _lib as L
    need = ctypes.c_int64(0)
    M._raise(L.load().mfla_refine_workspace_size(N, _REFINE_KIND[kind] | flag, n, nrhs, ctypes.byref(need)))
    key = (torch.cuda.current_device(), N, kind, flag)
    w = _ws_cache.get(key)
    if w is None or w.numel() < need.value:
        w = torch.empty(max(need.value, 1), dtype=torch.float64, device="cuda")
        _ws_cache[key] = w
    return w, need.value


def refine_solve(A, B, kind="cholesky", *, tol=None, max_iter=30, uplo="general", stages=False, residual="auto", check=True):
    """Solve A X = B (A, B Float64xN on the device, N = 2..4) with a binary64
    factor and Float64xN refinement, in the library's native driver
    (mfla_refine_solve). The default tolerance 2^-(52N+8) stops at the
    Float64xN residual floor (the plain Float64xN solve's backward error).
    Returns (X, history) -- history holds the binary64 backward-error
    estimate after each correction -- and with stages=True
    also milliseconds per stage (factor, solve, update, residual, check).
    residual="limbs" sums R -= A*D in as many limbs as |D|/|X| requires;
    residual="ozaki" uses the Ozaki-split INT8 matrix-core product (exact
    digit products recombined in Float64xN; n <= 8192); "auto" picks ozaki
    when n <= 8192. X is the best iterate; when its estimate is above tol
    (kappa(A) * u_64 >= 1, or tol below the Float64xN floor) check=True
    raises ConvergenceError (carrying X and the history), check=False
    returns it."""
    import ctypes
    from . import _lib as L
    N = M._limbs(A, B)
    if residual not in ("limbs", "ozaki", "auto"):
        raise M.ArgumentError("residual must be :auto, :limbs or :ozaki")
    if residual == "auto":
        residual = "ozaki" if A.shape[2] <= 8192 else "limbs"
    flag = 16 if residual == "ozaki" else 0
    tol = tol if tol is not None else 2.0 ** (-52 * N - 8)
    n, nrhs = A.shape[2], B.shape[1]
    X = torch.empty_like(B)
    work, length = _refine_workspace(N, kind, n, nrhs, flag)
    u = M._sym(uplo, M._UPLO_R, "uplo must be :general, :lower, or :upper")
    hist = np.zeros(max_iter)
    st = np.zeros(5)
    info = ctypes.c_int64(0)
    iters = ctypes.c_int(0)
    rc = L.load().mfla_refine_solve(M._handle(), N, _REFINE_KIND[kind] | flag, u, M._mat(A), M._mat(B), M._mat(X),
                                    ctypes.c_void_p(work.data_ptr()), length, float(tol), int(max_iter), ctypes.byref(info),
                                    ctypes.byref(iters), M._sp(hist), M._sp(st) if stages else None)
    h = [float(v) for v in hist[: iters.value]]
    if rc == L.MFLA_NOT_CONVERGED:
        if check:
            raise M.ConvergenceError(f"refinement stopped at backward error {min(h):.3e} > tol {tol:.3e}", X, h)
    else:
        M._raise(rc)
    if stages:
        return X, h, {k: round(float(v), 3) for k, v in zip(_STAGES, st)}
    return X, h


def refine_solve_pieces(A, B, kind="cholesky", *, block=64, tol=None, max_iter=30, uplo="general", stages=False):
    """The same refinement written as a caller would with MFLA's public
    pieces: factor the leading limb with cholesky!/lu!, then ldiv! on the
    residual's leading limb, add_mixed_, residual! and
    normwise_backward_error per iteration."""
    N = A.shape[0]
    tol = tol if tol is not None else 64.0 * 2.0 ** (-52 * N)
    T = _Timer(stages)
    T.mark("start")
    F64 = A[0:1].clone()
    if kind == "cholesky":
        F = M.cholesky_(F64, config=M.KernelConfig(cholesky_block=block))
    else:
        F = M.lu_(F64, config=M.KernelConfig(lu_block=block))
    X = torch.zeros_like(B)
    R = B.clone()
    T.mark("factor")
    hist = []
    for _ in range(max_iter):
        D = R[0:1]                                     # leading limb of the residual, in place
        M.ldiv_(D, F, config=M.KernelConfig(trsm_order="blocked"))
        T.mark("solve")
        M.add_mixed_(X, D)
        T.mark("update")
        M.residual_(R, A, X, B, uplo=uplo)
        T.mark("residual")
        e = _max_be(M.normwise_backward_error(A, X, B, R, uplo=uplo))
        T.mark("check")
        hist.append(e)
        if e <= tol or (len(hist) > 1 and e > 0.5 * hist[-2]):
            break
    if stages:
        return X, hist, T.done()
    return X, hist


def plain_solve(A, B, kind="cholesky", *, block=16):
    """The plain Float64xN factor + solve (MFLA's own cholesky!/lu! + ldiv!)."""
    F0 = A.clone()
    if kind == "cholesky":
        F = M.cholesky_(F0, config=M.KernelConfig(cholesky_block=block))
    else:
        F = M.lu_(F0, config=M.KernelConfig(lu_block=block))
    return M.solve(F, B, config=M.KernelConfig(trsm_order="blocked"))


def bench_refine(n=4096, nrhs=64, kind="cholesky", kappa=1e8, N=4, seed=20260812, residual="auto", keep_cols=None):
    """Times the refined and the plain Float64xN solve on the same matrix and
    right-hand sides; reports times and the Float64xN-evaluated backward
    errors. keep_cols: right-hand-side columns whose solutions (both
    solves), right-hand sides and the matrix are returned as host arrays
    under "_check", for an exact evaluation of the residual by the caller."""
    A = spd_kappa(n, kappa, N=N, seed=seed) if kind == "cholesky" else general_kappa(n, kappa, N=N, seed=seed)
    rng = np.random.default_rng(seed + 1)
    Bh = np.zeros((N, nrhs, n))
    Bh[0] = rng.standard_normal((nrhs, n))
    for l in range(1, N):
        Bh[l] = Bh[l - 1] * 2.0**-53 * rng.uniform(-1, 1, (nrhs, n))
    B = torch.from_numpy(Bh).cuda()
    out = {"n": n, "nrhs": nrhs, "kind": kind, "kappa": kappa, "limbs": N, "residual": residual,
           "orthogonal_factors": _FACTORS[ORTHOGONAL]}
    keep = {}
    for name, fn in (("refined", lambda: refine_solve(A, B, kind, residual=residual)), ("plain", lambda: (plain_solve(A, B, kind), None))):
        fn()                                   # warm-up: code-object loading, workspace allocation
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        X, hist = fn()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        R = torch.zeros_like(B)
        M.residual_(R, A, X, B)
        be = _max_be(M.normwise_backward_error(A, X, B, R))
        out[name] = {"seconds": round(dt, 4), "backward_error": be}
        if keep_cols is not None:
            keep[name] = X[:, keep_cols, :].cpu().numpy()
        if hist is not None:
            out[name]["iterations"] = len(hist)
            out[name]["history"] = [float(f"{h:.3e}") for h in hist]
    out["speedup"] = round(out["plain"]["seconds"] / out["refined"]["seconds"], 3)
    _, _, st = refine_solve(A, B, kind, stages=True, residual=residual)
    out["refined"]["stage_ms"] = st
    P = _Timer(True)
    P.mark("start")
    F0 = A.clone()
    F = M.cholesky_(F0) if kind == "cholesky" else M.lu_(F0)
    P.mark("factor")
    M.solve(F, B, config=M.KernelConfig(trsm_order="blocked"))
    P.mark("solve")
    out["plain"]["stage_ms"] = P.done()
    if keep_cols is not None:
        keep["A"] = A.cpu().numpy()
        keep["B"] = np.ascontiguousarray(Bh[:, keep_cols, :])
        keep["cols"] = list(keep_cols)
        out["_check"] = keep
    return out
