"""Reverse-mode rules for solve, Cholesky and logdet on Float64xN matrices.

MFLA has no differentiation support; these are the pullbacks its Julia
callers would register as ChainRulesCore rrules, here as
torch.autograd.Function classes over limb-planar tensors. Each pullback
reuses the Cholesky factor saved by the forward pass (no refactoring) and
runs the library's adjoint kernels:

    x = S^{-1} B      Bbar = S^{-1} Xbar,   Sbar = -sym(Bbar X^T)
    y = logdet(S)     Sbar = ybar S^{-1}
    L = chol(S)       Sbar = sym(L^{-T} Phi(L^T Lbar) L^{-1})

sym(G) = (G + G^T)/2: gradients with respect to a symmetric S are returned
symmetric. Tensors are (N, cols, rows) limb planes; a scalar is an (N,)
tensor of limbs. Gradients are Float64xN values in the same layout, so
torch must not add two of them (limb-wise sums are not Float64xN sums): use
each input once per graph, or combine gradients with mfla.api.arith("add").
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib as L
from . import api as M


def _chol(S, config=None):
    F = S.clone()
    f = M.cholesky_(F, config=config)
    return f


def logdet(F: M.MFCholesky):
    """logdet(S) = 2 sum log L_ii from a Cholesky factor, as N host limbs."""
    N = F.factors.shape[0]
    out = np.zeros(N)
    M._raise(L.load().mfla_logdet(M._handle(), N, M._mat(F.factors), F.info, M._sp(out)))
    return out


def solve_adjoint(F: M.MFCholesky, X, Xbar, *, config=None):
    """(Bbar, Sbar) of x = S^{-1} B for the cotangent Xbar."""
    N = M._limbs(F.factors, X, Xbar)
    n = F.factors.shape[2]
    Bbar = torch.empty_like(Xbar)
    Sbar = torch.empty((N, n, n), dtype=torch.float64, device=X.device)
    M._raise(L.load().mfla_solve_adjoint(M._handle(), N, M._mat(F.factors), F.info, M._mat(X), M._mat(Xbar), M._mat(Bbar),
                                         M._mat(Sbar), M._cfg(config)))
    return Bbar, Sbar


def logdet_adjoint(F: M.MFCholesky, ybar, *, config=None):
    """Sbar = ybar * S^{-1} (ybar: N limbs)."""
    N = F.factors.shape[0]
    n = F.factors.shape[2]
    Sbar = torch.empty((N, n, n), dtype=torch.float64, device=F.factors.device)
    yb = M._scalar(np.asarray(ybar, dtype=np.float64), N)
    M._raise(L.load().mfla_logdet_adjoint(M._handle(), N, M._mat(F.factors), F.info, M._sp(yb), M._mat(Sbar), M._cfg(config)))
    return Sbar


def cholesky_adjoint(F: M.MFCholesky, Lbar, *, config=None):
    """Sbar of L = chol(S) for the cotangent Lbar (its lower triangle)."""
    N = M._limbs(F.factors, Lbar)
    Sbar = torch.empty_like(Lbar)
    work = torch.empty_like(Lbar)
    M._raise(L.load().mfla_cholesky_adjoint(M._handle(), N, M._mat(F.factors), F.info, M._mat(Lbar), M._mat(Sbar), M._mat(work),
                                            M._cfg(config)))
    return Sbar


def _lower(F):
    """The factor with its upper storage (MFLA leaves S's upper triangle
    there) cleared."""
    N, n, _ = F.shape
    idx = torch.arange(n, device=F.device)
    keep = (idx[None, :] >= idx[:, None])            # planar [col, row]: row >= col
    return torch.where(keep[None], F, torch.zeros((), dtype=F.dtype, device=F.device))


class CholeskySolve(torch.autograd.Function):
    """X = S^{-1} B through cholesky! + ldiv!; pullback reuses the factor."""

    @staticmethod
    def forward(ctx, S, B):
        F = _chol(S)
        X = M.solve(F, B)
        ctx.save_for_backward(F.factors, X)
        ctx.info = F.info
        return X

    @staticmethod
    def backward(ctx, Xbar):
        Fm, X = ctx.saved_tensors
        Bbar, Sbar = solve_adjoint(M.MFCholesky(Fm, ctx.info), X, Xbar.contiguous())
        return Sbar, Bbar


class LogDet(torch.autograd.Function):
    """y = logdet(S) (an (N,) tensor of limbs) through cholesky!."""

    @staticmethod
    def forward(ctx, S):
        F = _chol(S)
        ctx.save_for_backward(F.factors)
        ctx.info = F.info
        return torch.from_numpy(logdet(F)).to(S.device)

    @staticmethod
    def backward(ctx, ybar):
        (Fm,) = ctx.saved_tensors
        return logdet_adjoint(M.MFCholesky(Fm, ctx.info), ybar.detach().cpu().numpy())


class Cholesky(torch.autograd.Function):
    """L = chol(S) (lower triangle, zeros above) through cholesky!."""

    @staticmethod
    def forward(ctx, S):
        F = _chol(S)
        ctx.save_for_backward(F.factors)
        ctx.info = F.info
        return _lower(F.factors)

    @staticmethod
    def backward(ctx, Lbar):
        (Fm,) = ctx.saved_tensors
        return cholesky_adjoint(M.MFCholesky(Fm, ctx.info), Lbar.contiguous())


def cholesky_solve(S, B):
    return CholeskySolve.apply(S, B)


def logdet_fn(S):
    return LogDet.apply(S)


def cholesky_fn(S):
    return Cholesky.apply(S)
