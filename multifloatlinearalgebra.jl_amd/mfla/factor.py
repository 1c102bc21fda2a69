"""Factor protocol and diagnostics (src/types.jl:345-403, src/diagnostics.jl,
src/factorizations/ldlt.jl:20-63, src/factorizations/qr.jl:863-888).

Diagnostics are O(n) scans of factor metadata: the diagonal (n elements) is
copied to the host and combined with the library's own Float64xN networks
(``mfla_host_arith``); O(n^2) scans (max |U|, finiteness) run on the GPU.
Multi-double values are returned as numpy arrays of limbs.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib as L
from .api import (MFLU, MFLDLT, MFQR, AbstractMFFactorization, ArgumentError, MFCholesky, _handle, _mat, _raise, host_arith, rows_cols)


class _H:
    """Scalar Float64xN host arithmetic on (N,) limb arrays."""

    def __init__(self, N):
        self.N = N

    def _op(self, op, a, b=None):
        r = host_arith(op, np.asarray(a, dtype=np.float64).reshape(self.N, 1),
                       None if b is None else np.asarray(b, dtype=np.float64).reshape(self.N, 1))
        return r[:, 0]

    def c(self, v):
        x = np.zeros(self.N)
        x[0] = v
        return x

    def abs(self, a): return self._op("abs", a)
    def max(self, a, b): return self._op("max", a, b)
    def min(self, a, b): return self._op("min", a, b)
    def div(self, a, b): return self._op("div", a, b)
    def mul(self, a, b): return self._op("mul", a, b)
    def add(self, a, b): return self._op("add", a, b)
    def sub(self, a, b): return self._op("sub", a, b)
    def sqrt(self, a): return self._op("sqrt", a)
    def lt(self, a, b): return bool(self._op("lt", a, b)[0] != 0.0)
    def gt(self, a, b): return self.lt(b, a)

    @staticmethod
    def iszero(a): return bool(np.all(np.asarray(a) == 0.0))


# ------------------------------------------------------------- accessors
def factor_status(F): return F.info
def factor_kind(F): return F.kind
def factor_matrix(F): return F.factors
def issuccess(F): return F.info == 0
def factor_precision(F): return f"Float64x{F.factors.shape[0]}"
def factor_provider(F): return "mfla_hip"


def factor_state(F):
    """types.jl:394-403."""
    s = F.info
    if s == 0:
        return "success"
    if s == -1:
        return "nonfinite_input"
    if s < 0:
        return "numerical_breakdown"
    if F.kind == "cholesky":
        return "not_posdef"
    if F.kind in ("lu", "ldlt"):
        return "singular"
    return "numerical_breakdown"


def factor_pivots(F: MFLDLT): return F.pivots.cpu().numpy().copy()
def factor_blocks(F: MFLDLT): return F.blocks.cpu().numpy().copy()


def factor_rdiag(F: MFQR):
    """qr.jl:73-85: copy of the signed diagonal of R, shape (N, min(m, n))."""
    return _diag(F)


def factor_permutation(F):
    """ldlt.jl:41-63: p with A_original[p, p] = L*D*L' (one-based); for a QR
    factor (qr.jl:63-71) the column permutation with A_original[:, p] = Q R."""
    if isinstance(F, MFQR):
        return F.permutation.cpu().numpy().copy()
    blocks = factor_blocks(F)
    piv = factor_pivots(F)
    n = len(blocks)
    perm = np.arange(1, n + 1)
    k = 0
    while k < n:
        b = blocks[k]
        if b == 1:
            p = piv[k] - 1
            perm[k], perm[p] = perm[p], perm[k]
            k += 1
        elif b == 2 and k < n - 1:
            p = piv[k] - 1
            perm[k + 1], perm[p] = perm[p], perm[k + 1]
            k += 2
        elif b == 0:
            break
        else:
            raise ArgumentError("invalid LDLT block structure")
    return perm


def _diag(F, count=None):
    A = F.factors
    n = min(rows_cols(A))
    d = torch.diagonal(A, dim1=1, dim2=2)[:, :n] if A.dim() == 3 else A
    return d.cpu().numpy().copy()


def _absmax(A, region):
    out = np.zeros(4)
    N = A.shape[0]
    _raise(L.load().mfla_absmax(_handle(), N, region, _mat(A), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out[:N].copy()


def _finite(A, region):
    out = ctypes.c_int(0)
    _raise(L.load().mfla_all_finite(_handle(), A.shape[0], region, _mat(A), ctypes.byref(out)))
    return bool(out.value)


def _magnitude_range(h, diag, count):
    """diagnostics.jl:20-30."""
    if count <= 0:
        return None, None
    mn = h.abs(diag[:, 0])
    mx = mn.copy()
    for i in range(1, count):
        v = h.abs(diag[:, i])
        mn = h.min(mn, v)
        mx = h.max(mx, v)
    return mn, mx


def _optional_spread(h, mn, mx):
    """diagnostics.jl:14-18."""
    if mn is None:
        return None
    if h.iszero(mn):
        return h.c(np.inf)
    return h.div(mx, mn)


def _inertia_1x1(h, v):
    z = h.c(0.0)
    if h.gt(v, z):
        return (1, 0, 0)
    if h.lt(v, z):
        return (0, 1, 0)
    return (0, 0, 1)


def _scaled_block(h, d11, d21, d22):
    scale = h.max(h.max(h.abs(d11), h.abs(d21)), h.abs(d22))
    return scale, h.div(d11, scale), h.div(d21, scale), h.div(d22, scale)


def _inertia_2x2(h, d11, d21, d22):
    """diagnostics.jl:118-134."""
    scale = h.max(h.max(h.abs(d11), h.abs(d21)), h.abs(d22))
    if h.iszero(scale):
        return (0, 0, 2)
    a, b, c = h.div(d11, scale), h.div(d21, scale), h.div(d22, scale)
    det = h.sub(h.mul(a, c), h.mul(b, b))
    z = h.c(0.0)
    if h.lt(det, z):
        return (1, 1, 0)
    trace = h.add(a, c)
    if h.gt(det, z):
        return (2, 0, 0) if h.gt(trace, z) else (0, 2, 0)
    return (1, 0, 1) if h.gt(trace, z) else ((0, 1, 1) if h.lt(trace, z) else (0, 0, 2))


def _smallest_abs_eigenvalue(h, d11, d21, d22):
    """diagnostics.jl:136-153."""
    scale = h.max(h.max(h.abs(d11), h.abs(d21)), h.abs(d22))
    if h.iszero(scale):
        return h.c(0.0)
    a, b, c = h.div(d11, scale), h.div(d21, scale), h.div(d22, scale)
    amc = h.sub(a, c)
    disc = h.sqrt(h.add(h.mul(amc, amc), h.mul(h.mul(h.c(4.0), b), b)))
    apc = h.add(a, c)
    first = h.div(h.add(apc, disc), h.c(2.0))
    second = h.div(h.sub(apc, disc), h.c(2.0))
    largest = h.max(h.abs(first), h.abs(second))
    if h.iszero(largest):
        return h.c(0.0)
    det = h.abs(h.sub(h.mul(a, c), h.mul(b, b)))
    return h.mul(scale, h.div(det, largest))


def factor_inertia(F: MFLDLT):
    """diagnostics.jl:162-185."""
    h = _H(F.factors.shape[0])
    blocks = factor_blocks(F)
    diag = _diag(F)
    dsub = F.dsub.cpu().numpy()
    pos = neg = zer = 0
    k, n = 0, len(blocks)
    while k < n:
        b = blocks[k]
        if b == 1:
            p, q, z = _inertia_1x1(h, diag[:, k])
            k += 1
        elif b == 2 and k < n - 1:
            p, q, z = _inertia_2x2(h, diag[:, k], dsub[:, k], diag[:, k + 1])
            k += 2
        else:
            break
        pos += p
        neg += q
        zer += z
    return {"positive": pos, "negative": neg, "zero": zer}


def factor_diagnostics(F: AbstractMFFactorization, *, atol=0.0, rtol=0.0):
    """diagnostics.jl:32-275 (atol / rtol: the QR rank thresholds)."""
    N = F.factors.shape[0]
    h = _H(N)
    base = {
        "kind": F.kind, "status": F.info, "state": factor_state(F), "success": issuccess(F),
        "precision": factor_precision(F), "provider": factor_provider(F),
        "failure_location": F.info if F.info > 0 else None,
    }
    if isinstance(F, MFQR):
        diag = _diag(F)
        k = diag.shape[1]
        mn, mx = _magnitude_range(h, diag, k) if k else (None, None)
        tau_ok = bool(torch.isfinite(F.tau).all().item())
        base.update(failure_location=None, permutation=factor_permutation(F), rdiag=diag, minimum_rdiag=mn, maximum_rdiag=mx,
                    rdiag_spread=_optional_spread(h, mn, mx),
                    rank_at_threshold=numerical_rank(F, atol=atol, rtol=rtol) if issuccess(F) else 0,
                    atol=atol, rtol=rtol, finite=_finite(F.factors, 0) and tau_ok)
        return base
    if isinstance(F, MFCholesky):
        n = rows_cols(F.factors)[0]
        accepted = n if F.info == 0 else (F.info - 1 if F.info > 0 else 0)
        mn, mx = _magnitude_range(h, _diag(F), accepted)
        base.update(accepted_pivots=accepted, minimum_diagonal=mn, maximum_diagonal=mx,
                    diagonal_spread=_optional_spread(h, mn, mx), finite=_finite(F.factors, 1))
        return base
    if isinstance(F, MFLU):
        npiv = F.ipiv.shape[0]
        accepted = F.info - 1 if F.info > 0 else (npiv if F.info == 0 else 0)
        mn, mx = _magnitude_range(h, _diag(F), accepted)
        max_u = _absmax(F.factors, 2)
        om = F.original_maximum
        growth = None if h.iszero(om) else h.div(max_u, om)
        base.update(accepted_pivots=accepted, pivots=F.ipiv.cpu().numpy().copy(), minimum_pivot=mn, maximum_pivot=mx,
                    original_maximum=None if F.info == -1 else om, maximum_u=max_u, pivot_growth=growth,
                    finite=_finite(F.factors, 0))
        return base
    if isinstance(F, MFLDLT):
        blocks = factor_blocks(F)
        diag = _diag(F)
        dsub = F.dsub.cpu().numpy()
        one = two = 0
        min_block = None
        max_entry = h.c(0.0)
        k, n = 0, len(blocks)
        while k < n:
            b = blocks[k]
            if b == 1:
                one += 1
                v = h.abs(diag[:, k])
                min_block = v if min_block is None else h.min(min_block, v)
                max_entry = h.max(max_entry, v)
                k += 1
            elif b == 2 and k < n - 1:
                two += 1
                v = _smallest_abs_eigenvalue(h, diag[:, k], dsub[:, k], diag[:, k + 1])
                min_block = v if min_block is None else h.min(min_block, v)
                max_entry = h.max(h.max(h.max(max_entry, h.abs(diag[:, k])), h.abs(dsub[:, k])), h.abs(diag[:, k + 1]))
                k += 2
            else:
                break
        om = F.original_maximum
        min_scaled = None if (min_block is None or h.iszero(om)) else h.div(min_block, om)
        growth = None if h.iszero(om) else h.div(max_entry, om)
        base.update(one_by_one_pivots=one, two_by_two_pivots=two, pivots=factor_pivots(F), blocks=blocks,
                    inertia=factor_inertia(F), minimum_block_eigenvalue_magnitude=min_block, minimum_scaled_block=min_scaled,
                    original_maximum=None if F.info == -1 else om, maximum_block_entry=max_entry, block_growth=growth,
                    finite=_finite(F.factors, 0) and bool(np.all(np.isfinite(dsub))))
        return base
    raise ArgumentError("unsupported factorization")


def numerical_rank(F, atol=0.0, rtol=0.0):
    """numerical_rank (src/factorizations/qr.jl:863-888): leading count of
    diagonal entries with |R[i,i]| > max(atol, rtol * max|R[i,i]|). MFLA
    defines it for its QR factor; here it scans the triangular factor's
    diagonal (U of LU, L of Cholesky)."""
    if not issuccess(F):
        raise ArgumentError("numerical_rank requires a successful factorization")
    N = F.factors.shape[0]
    h = _H(N)
    at, rt = np.zeros(N), np.zeros(N)
    at[: len(np.atleast_1d(atol))] = np.atleast_1d(atol)[:N]
    rt[: len(np.atleast_1d(rtol))] = np.atleast_1d(rtol)[:N]
    if not (np.all(np.isfinite(at)) and not h.lt(at, h.c(0.0))):
        raise ArgumentError("atol must be finite and nonnegative")
    if not (np.all(np.isfinite(rt)) and not h.lt(rt, h.c(0.0))):
        raise ArgumentError("rtol must be finite and nonnegative")
    diag = _diag(F)
    count = diag.shape[1]
    largest = h.c(0.0)
    for i in range(count):
        largest = h.max(largest, h.abs(diag[:, i]))
    thr = h.max(at, h.mul(rt, largest))
    rank = 0
    for i in range(count):
        if not h.gt(h.abs(diag[:, i]), thr):
            break
        rank += 1
    return rank
