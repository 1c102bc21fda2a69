"""Python mirror of MultiFloatLinearAlgebra.jl's public API over the HIP library.

Names follow the reference with Julia's ``!`` spelled ``_`` (``gemm!`` ->
``gemm_``). Arguments keep MFLA's order and meaning, keyword symbols are
strings (``side="left"``), and errors raise the exception MFLA throws
(DimensionMismatch, ArgumentError, PosDefException, SingularException,
DomainError) before any memory is touched.

Array convention: a Float64xN matrix of shape (m, n) is a CUDA float64 torch
tensor of shape (N, n, m) whose last dimension is contiguous -- N limb planes,
each column-major. A vector of length n is a tensor of shape (N, n).
Sub-matrix views (``A[:, j0:j1, i0:i1]``) are accepted without copies.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import torch

from . import _lib as L


# --------------------------------------------------------------- exceptions
class DimensionMismatch(ValueError):
    pass


class ArgumentError(ValueError):
    pass


class DomainError(ValueError):
    pass


class BoundsError(IndexError):
    pass


class PosDefException(Exception):
    def __init__(self, info):
        super().__init__(f"PosDefException({info})")
        self.info = info


class SingularException(Exception):
    def __init__(self, info):
        super().__init__(f"SingularException({info})")
        self.info = info


class ConvergenceError(Exception):
    """Iterative refinement stopped above its tolerance. ``X`` is the best
    iterate (still a valid approximation) and ``history`` the estimates."""

    def __init__(self, msg, X=None, history=None):
        super().__init__(msg)
        self.X, self.history = X, history


def _raise(rc: int):
    if rc == L.MFLA_OK:
        return
    msg = L.last_error()
    if rc == L.MFLA_ERR_DIMENSION:
        raise DimensionMismatch(msg)
    if rc in (L.MFLA_ERR_ARGUMENT, L.MFLA_ERR_ALIAS):
        raise ArgumentError(msg)
    if rc == L.MFLA_ERR_BOUNDS:
        raise BoundsError(msg)
    if rc == L.MFLA_ERR_FAILED_FACTOR:
        info = int(msg[msg.find("(") + 1: msg.find(")")]) if "(" in msg else 0
        if msg.startswith("PosDef"):
            raise PosDefException(info)
        raise SingularException(info)
    raise RuntimeError(f"mfla error {rc}: {msg}")


# ------------------------------------------------------------------ handle
_handles: dict = {}


def _handle():
    dev = torch.cuda.current_device()
    h = _handles.get(dev)
    if h is None:
        lib = L.load()
        p = ctypes.c_void_p()
        _raise(lib.mfla_handle_create(ctypes.byref(p)))
        h = p
        _handles[dev] = h
    L.load().mfla_handle_set_stream(h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return h


def synchronize():
    _raise(L.load().mfla_synchronize(_handle()))


# ------------------------------------------------------------------ config
_GEMM = {"auto": 0, "direct": 1, "packed": 2, "fused": 3, "ozaki": 4}
_LDLT = {"auto": 0, "unblocked": 1, "blocked": 2}
_ORDER = {"auto": 0, "exact": 1, "blocked": 2}


@dataclass
class KernelConfig:
    """KernelConfig (src/types.jl:13-27) plus the device-only knobs
    ``trsm_order`` (backward-substitution order) and ``gemm_tile``."""

    reduction_tile: int = 64
    column_tile: int = 16
    cholesky_block: int = 16
    lu_block: int = 16
    ldlt_block: int = 0
    ldlt_strategy: str = "auto"
    ldlt_blocked_crossover: int = 512
    thread_count: int = 1
    gemm_strategy: str = "auto"
    gemm_packed_crossover: int = 2**63 - 1
    gemm_panel_columns: int = 0
    gemm_micro_columns: int = 0
    trsm_order: str = "auto"
    gemm_tile: int = -1

    def _c(self):
        if self.gemm_strategy not in _GEMM:
            raise ArgumentError("gemm_strategy must be :auto, :direct, :packed, :fused, or :ozaki")
        if self.ldlt_strategy not in _LDLT:
            raise ArgumentError("ldlt_strategy must be :auto, :unblocked, or :blocked")
        c = L.MFConfig()
        c.reduction_tile = self.reduction_tile
        c.column_tile = self.column_tile
        c.cholesky_block = self.cholesky_block
        c.lu_block = self.lu_block
        c.ldlt_block = self.ldlt_block
        c.ldlt_strategy = _LDLT[self.ldlt_strategy]
        c.gemm_strategy = _GEMM[self.gemm_strategy]
        c.ldlt_blocked_crossover = self.ldlt_blocked_crossover
        c.thread_count = self.thread_count
        c.gemm_panel_columns = self.gemm_panel_columns
        c.gemm_micro_columns = self.gemm_micro_columns
        c.trsm_order = _ORDER[self.trsm_order]
        c.gemm_tile = self.gemm_tile
        c.gemm_packed_crossover = self.gemm_packed_crossover
        return c


def _cfg(config):
    return ctypes.byref((config or KernelConfig())._c())


# ------------------------------------------------------------------- plans
_PLAN_REASON = {1: "forced_direct", 2: "forced_packed", 3: "forced_fused", 4: "forced_ozaki",
                5: "auto_outside_calibrated_shape", 6: "auto_reduction_too_small", 7: "auto_below_crossover",
                8: "auto_above_crossover", 9: "auto_fused_direct", 10: "forced_unblocked", 11: "forced_blocked"}
_GEMM_NAME = {v: k for k, v in _GEMM.items()}
_LDLT_NAME = {v: k for k, v in _LDLT.items()}


@dataclass(frozen=True)
class GemmPlan:
    """GemmPlan (src/types.jl:210-217): MFLA's route and CPU decomposition,
    plus the device launch gemm! makes for it (tile index and shape, network
    "mulacc_x3" or "standard", launches, workgroups, x3 remainder rows)."""

    strategy: str
    reason: str
    panel_columns: int
    micro_columns: int
    workers: int
    packed_elements_per_worker: int
    tile: int
    tile_rows: int
    tile_cols: int
    network: str
    launches: int
    workgroups: int
    remainder_rows: int


@dataclass(frozen=True)
class LDLTPlan:
    """LDLTPlan (src/types.jl:219-223)."""

    strategy: str
    reason: str
    block_size: int


def gemm_plan(limbs: int, m: int, k: int, n: int, config: KernelConfig | None = None) -> GemmPlan:
    """gemm_plan(T, m, k, n, config) (src/kernels/gemm.jl:44-136) for
    Float64x{limbs}: the same strategy / reason resolution (volume-based
    packed crossover on near-square shapes, fused mulacc_x3 for x3 on :auto)
    and the device launch that gemm! runs for it."""
    p = L.MFGemmPlan()
    _raise(L.load().mfla_gemm_plan(int(limbs), int(m), int(k), int(n), _cfg(config), ctypes.byref(p)))
    return GemmPlan(_GEMM_NAME[p.strategy], _PLAN_REASON[p.reason], p.panel_columns, p.micro_columns, p.workers,
                    p.packed_elements_per_worker, p.tile, p.tile_rows, p.tile_cols, "mulacc_x3" if p.network else "standard",
                    p.launches, p.workgroups, p.remainder_rows)


def ldlt_plan(limbs: int, n: int, config: KernelConfig | None = None) -> LDLTPlan:
    """ldlt_plan(T, n, config) (src/factorizations/ldlt.jl:329-350): the
    unblocked route below the crossover (512), the lazy blocked panel of
    width 16 / 12 / 8 above it; ldlt! follows this plan."""
    p = L.MFLDLTPlan()
    _raise(L.load().mfla_ldlt_plan(int(limbs), int(n), _cfg(config), ctypes.byref(p)))
    return LDLTPlan(_LDLT_NAME[p.strategy], _PLAN_REASON[p.reason], p.block_size)


# ------------------------------------------------------------------ arrays
def _limbs(*ts):
    N = ts[0].shape[0]
    for t in ts:
        if t.shape[0] != N:
            raise ArgumentError("all operands must have the same MultiFloat limb count")
    if not 1 <= N <= 4:
        raise ArgumentError("only 1-4 limb MultiFloats are supported")
    return N


def _mat(t: torch.Tensor) -> L.MFMatrix:
    """Descriptor of a (N, cols, rows) tensor (or (N, n) vector as n x 1)."""
    if not t.is_cuda or t.dtype != torch.float64:
        raise ArgumentError("operands must be CUDA float64 tensors")
    if t.dim() == 2:
        n = t.shape[1]
        if n > 1 and t.stride(1) != 1:
            raise ArgumentError("vector limbs must be contiguous")
        return L.MFMatrix(t.data_ptr(), n, 1, max(n, 1), t.stride(0))
    if t.dim() != 3:
        raise ArgumentError("operands must have shape (limbs, cols, rows)")
    N, cols, rows = t.shape
    if t.numel() == 0:
        return L.MFMatrix(t.data_ptr(), rows, cols, max(rows, 1), max(rows * cols, 1))
    if rows > 1 and t.stride(2) != 1:
        raise ArgumentError("matrix limb planes must be column-major (last dim contiguous)")
    ld = t.stride(1) if cols > 1 else max(rows, 1)
    return L.MFMatrix(t.data_ptr(), rows, cols, max(ld, 1), t.stride(0))


def _scalar(v, N):
    out = np.zeros(4)
    a = np.asarray(v, dtype=np.float64).reshape(-1)
    out[: min(len(a), N)] = a[:N]
    return out


def _sp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def rows_cols(t):
    return (t.shape[1], 1) if t.dim() == 2 else (t.shape[2], t.shape[1])


# ----------------------------------------------------------------- BLAS 3
def gemm_(C, A, B, alpha=1.0, beta=0.0, *, config: Optional[KernelConfig] = None, workspace=None):
    """gemm! (src/kernels/gemm.jl:747-774): C = alpha*A*B + beta*C."""
    N = _limbs(C, A, B)
    _raise(L.load().mfla_gemm(_handle(), N, _mat(C), _mat(A), _mat(B), _sp(_scalar(alpha, N)), _sp(_scalar(beta, N)), _cfg(config)))
    return C


def syrk_(C, P, alpha=1.0, beta=0.0, *, uplo="lower", config=None):
    """syrk! (src/kernels/syrk.jl:67-108): lower C = alpha*P'*P + beta*C."""
    N = _limbs(C, P)
    u = 0 if uplo in ("lower", "L") else 1
    _raise(L.load().mfla_syrk(_handle(), N, u, _mat(C), _mat(P), _sp(_scalar(alpha, N)), _sp(_scalar(beta, N)), _cfg(config)))
    return C


def gemmt_(C, Lm, R, alpha=1.0, beta=0.0, *, config=None):
    """gemmt! (src/kernels/gemmt.jl:69-108): lower C = alpha*L*R' + beta*C."""
    N = _limbs(C, Lm, R)
    _raise(L.load().mfla_gemmt(_handle(), N, _mat(C), _mat(Lm), _mat(R), _sp(_scalar(alpha, N)), _sp(_scalar(beta, N)), _cfg(config)))
    return C


def _sym(v, table, what):
    if v not in table:
        raise ArgumentError(what)
    return table[v]


def trsm_(B, A, alpha=1.0, *, side="left", uplo="lower", trans="N", diag="nonunit", config=None):
    """trsm! (src/kernels/trsm.jl:224-302), in place on B."""
    N = _limbs(B, A)
    s = _sym(side, {"left": 0, "right": 1}, "side must be :left or :right")
    u = _sym(uplo, {"lower": 0, "upper": 1}, "uplo must be :lower or :upper")
    t = _sym(trans, {"N": 0, "T": 1}, "trans must be :N or :T")
    d = _sym(diag, {"nonunit": 0, "unit": 1}, "diag must be :unit or :nonunit")
    _raise(L.load().mfla_trsm(_handle(), N, s, u, t, d, _mat(B), _mat(A), _sp(_scalar(alpha, N)), _cfg(config)))
    return B


def trmm_(B, A, alpha=1.0, *, side="left", uplo="lower", trans="N", diag="nonunit", config=None):
    """trmm! (src/kernels/trmm.jl:75-131): B = alpha*op(A)*B or alpha*B*op(A)."""
    N = _limbs(B, A)
    sd = _sym(side, {"left": 0, "right": 1}, "side must be :left or :right")
    up = _sym(uplo, {"lower": 0, "upper": 1}, "uplo must be :lower or :upper")
    tr = _sym(trans, {"N": 0, "T": 1}, "trans must be :N or :T")
    dg = _sym(diag, {"nonunit": 0, "unit": 1}, "diag must be :unit or :nonunit")
    _raise(L.load().mfla_trmm(_handle(), N, sd, up, tr, dg, _mat(B), _mat(A), _sp(_scalar(alpha, N)), _cfg(config)))
    return B


def mfdot(x, y):
    """mfdot (src/kernels/dot.jl:1-22): one ascending chain over the
    column-major linear index. Returns the N limbs (numpy)."""
    N = _limbs(x, y)
    if x.shape != y.shape:
        raise DimensionMismatch("dot product shapes differ")
    out = np.zeros(N)
    _raise(L.load().mfla_dot(_handle(), N, _mat(x), _mat(y), _sp(out)))
    return out


def syrk_packed_(packed, P, alpha=1.0, beta=0.0, *, reduction_first=1, reduction_last=None, config=None):
    """syrk_packed! (src/kernels/syrk.jl:181-254): packed is an (N, n(n+1)/2)
    tensor holding the column-major lower triangle."""
    N = _limbs(packed, P)
    rl = P.shape[2] if reduction_last is None else reduction_last
    _raise(L.load().mfla_syrk_packed(_handle(), N, _mat(packed), _mat(P), _sp(_scalar(alpha, N)), _sp(_scalar(beta, N)),
                                     int(reduction_first), int(rl), _cfg(config)))
    return packed


def capabilities(N):
    """capabilities(Float64xN) (src/capabilities.jl:15-74) for this backend:
    the same facts with provider "mfla_hip", plus the device extensions."""
    supported = 1 <= N <= 4
    targets = {"x2": False, "x3": supported and N == 2, "x4": supported and N in (2, 3)}
    ops = ("dot", "gemv", "transpose_gemv", "symv", "gemm", "gemmt", "syrk", "syrk_packed", "trsv", "trsm", "trmm",
           "cholesky", "lu", "ldlt", "ldlt_lightweight_metadata", "factor_diagnostics", "vector_rhs", "multi_rhs",
           "residual", "refinement_correction", "reusable_workspace")
    caps = {"provider": "mfla_hip", "scalar_type": f"Float64x{N}", "base_type": "Float64", "supported": supported,
            "limb_count": N}
    caps.update({op: supported for op in ops})
    caps.update({
        "rrqr": False, "apply_q": False, "solve_r": False,
        "mixed_precision_residual": any(targets.values()), "mixed_residual_targets": targets,
        "mixed_residual_target_types": tuple(f"Float64x{k}" for k in (3, 4) if targets[f"x{k}"]),
        "factor_metadata_ownership": "factor_owned", "factor_matrix_ownership": "borrowed_input",
        "factorization_destructive": True, "factor_solve_mutates_factor": False,
        "shared_gemm_workspace_concurrency": "serialized_safe", "concurrent_factor_workspace": False,
        "syrk_authoritative_triangle": "lower", "syrk_inactive_triangle": "preserved",
        "threading": False, "stream_ordered": supported,
        "batched_solve": supported, "binary64_refinement": supported and N >= 2, "reverse_mode_rules": supported,
    })
    return caps


# ----------------------------------------------------------------- BLAS 2
def trsv_(x, A, alpha=1.0, *, uplo="lower", trans="N", diag="nonunit", config=None):
    """trsv! (src/kernels/trsv.jl:19-73), in place on the vector x."""
    N = _limbs(x, A)
    u = _sym(uplo, {"lower": 0, "upper": 1}, "uplo must be :lower or :upper")
    t = _sym(trans, {"N": 0, "T": 1}, "trans must be :N or :T")
    d = _sym(diag, {"nonunit": 0, "unit": 1}, "diag must be :unit or :nonunit")
    _raise(L.load().mfla_trsv(_handle(), N, u, t, d, _mat(x), _mat(A), _sp(_scalar(alpha, N)), _cfg(config)))
    return x


def gemv_(y, A, x, alpha=1.0, beta=0.0, *, trans="N", config=None):
    """gemv! (src/kernels/gemv.jl:111-167)."""
    N = _limbs(y, A, x)
    t = _sym(trans, {"N": 0, "T": 1}, "trans must be :N or :T")
    _raise(L.load().mfla_gemv(_handle(), N, t, _mat(y), _mat(A), _mat(x), _sp(_scalar(alpha, N)), _sp(_scalar(beta, N)), _cfg(config)))
    return y


def symv_(y, A, x, alpha=1.0, beta=0.0, *, uplo="lower", config=None):
    """symv! (src/kernels/symv.jl:53-89)."""
    N = _limbs(y, A, x)
    u = _sym(uplo, {"lower": 0, "upper": 1}, "uplo must be :lower or :upper")
    _raise(L.load().mfla_symv(_handle(), N, u, _mat(y), _mat(A), _mat(x), _sp(_scalar(alpha, N)), _sp(_scalar(beta, N)), _cfg(config)))
    return y


# -------------------------------------------------------------- workspace
class MFWorkspace:
    """MFWorkspace (src/types.jl:71-123): caller-owned device scratch."""

    def __init__(self, limbs: int, factor_capacity: int = 0, ldlt_block_capacity: int = 0):
        p = ctypes.c_void_p()
        _raise(L.load().mfla_workspace_create(_handle(), limbs, factor_capacity, ldlt_block_capacity, ctypes.byref(p)))
        self._p = p
        self.limbs = limbs

    def __del__(self):
        try:
            if self._p:
                L.load().mfla_workspace_destroy(self._p)
        except Exception:
            pass


def workspace_capacity(ws: MFWorkspace):
    out = (ctypes.c_int64 * 2)()
    _raise(L.load().mfla_workspace_capacity(ws._p, out))
    return {"factor": int(out[0]), "ldlt_block": int(out[1])}


def ensure_workspace_capacity_(ws: MFWorkspace, factor_capacity=0, ldlt_block_capacity=0):
    _raise(L.load().mfla_workspace_ensure(ws._p, factor_capacity, ldlt_block_capacity))
    return ws


# ---------------------------------------------------------- factorizations
class AbstractMFFactorization:
    kind = None


@dataclass
class MFCholesky(AbstractMFFactorization):
    factors: torch.Tensor
    info: int
    kind = "cholesky"


@dataclass
class MFLU(AbstractMFFactorization):
    factors: torch.Tensor
    ipiv: torch.Tensor
    info: int
    original_maximum: np.ndarray
    kind = "lu"


@dataclass
class MFLDLT(AbstractMFFactorization):
    factors: torch.Tensor
    dsub: torch.Tensor
    pivots: torch.Tensor
    blocks: torch.Tensor
    info: int
    original_maximum: np.ndarray
    kind = "ldlt"


@dataclass
class MFQR(AbstractMFFactorization):
    factors: torch.Tensor
    tau: torch.Tensor              # (N, min(m, n)) Householder scalars
    permutation: torch.Tensor      # int64, one-based: A_original[:, p] = Q R
    permutation_cycle_leaders: list
    info: int
    kind = "qr"


def _cycle_leaders(perm):
    """qr.jl:43-61: first index of every permutation cycle longer than one."""
    p = [int(v) for v in perm]
    seen = [False] * len(p)
    leaders = []
    for start in range(len(p)):
        if seen[start]:
            continue
        cur, length = start, 0
        while not seen[cur]:
            seen[cur] = True
            cur = p[cur] - 1
            length += 1
        if length > 1:
            leaders.append(start + 1)
    return leaders


def _qr_finish(A, tau, perm, info, check, name):
    if check and info == -1:
        raise DomainError(f"{name}: input matrix contains non-finite entries")
    return MFQR(A, tau, perm, _cycle_leaders(perm.cpu().numpy()), info)


_QR_ROUTE = {"auto": 0, "unblocked": 1, "blocked": 2}


def rrqr_(A, *, check=True, workspace=None, route="auto"):
    """rrqr! (src/factorizations/qr.jl:692-743): deterministic column-pivoted
    Householder QR in place; A_original[:, p] = Q R with p =
    factor_permutation(F). route: "auto" (MFLA's choice: the blocked DLAQPS
    route when min(m, n) >= 16 and m n >= 16384), "unblocked" or "blocked".
    The workspace argument is accepted for API parity (results do not depend
    on it)."""
    N = _limbs(A)
    m, n = rows_cols(A)
    r = _sym(route, _QR_ROUTE, "route must be :auto, :unblocked, or :blocked")
    tau = torch.zeros((N, min(m, n)), dtype=torch.float64, device=A.device)
    perm = torch.arange(1, n + 1, dtype=torch.int64, device=A.device)
    info = ctypes.c_int64(0)
    _raise(L.load().mfla_geqp3(_handle(), N, _mat(A), _mat(tau), ctypes.c_void_p(perm.data_ptr()), r, ctypes.byref(info)))
    return _qr_finish(A, tau, perm, int(info.value), check, "rrqr!")


def qr_(A, permutation=None, *, check=True, workspace=None):
    """qr! (src/factorizations/qr.jl:763-853): Householder QR without
    pivoting, or in the caller's column order (permutation[i] = original
    column of factor column i; the columns of A are reordered in place)."""
    N = _limbs(A)
    m, n = rows_cols(A)
    if permutation is not None:
        perm_h = np.asarray([int(v) for v in permutation], dtype=np.int64)
        if perm_h.shape[0] != n:
            raise DimensionMismatch(f"permutation length {perm_h.shape[0]} does not match {n} columns")
        seen = np.zeros(n, dtype=bool)
        for c in perm_h:
            if not 1 <= c <= n:
                raise ArgumentError(f"permutation entry {c} is outside 1:{n}")
            if seen[c - 1]:
                raise ArgumentError(f"permutation repeats column {c}")
            seen[c - 1] = True
        idx = torch.from_numpy(perm_h - 1).to(A.device)
        A.copy_(A.index_select(1, idx))
    tau = torch.zeros((N, min(m, n)), dtype=torch.float64, device=A.device)
    info = ctypes.c_int64(0)
    _raise(L.load().mfla_geqrf(_handle(), N, _mat(A), _mat(tau), ctypes.byref(info)))
    if permutation is not None:
        perm = torch.from_numpy(perm_h).to(A.device)
    else:
        perm = torch.arange(1, n + 1, dtype=torch.int64, device=A.device)
    return _qr_finish(A, tau, perm, int(info.value), check, "qr!")


def apply_q_(B, F, *, trans="N"):
    """apply_q! (src/factorizations/qr.jl:938-956): B <- Q B (trans = "N") or
    Q' B ("T") from the compact reflectors; B is a vector (N, m) or a matrix."""
    if not isinstance(F, MFQR) or F.info != 0:
        raise ArgumentError("apply_q! requires a successful QR factorization")
    t = _sym(trans, {"N": 0, "T": 1}, "trans must be :N or :T")
    N = _limbs(B, F.factors)
    X = _as_matrix(B)
    _raise(L.load().mfla_ormqr(_handle(), N, t, _mat(F.factors), _mat(F.tau), _mat(X)))
    return B


def solve_r_(B, F, rank, *, trans="N", config=None):
    """solve_r! (src/factorizations/qr.jl:958-1014): solve the leading
    rank x rank block of R (or its transpose) in place; B has rank rows."""
    if not isinstance(F, MFQR) or F.info != 0:
        raise ArgumentError("solve_r! requires a successful QR factorization")
    if trans not in ("N", "T"):
        raise ArgumentError("trans must be :N or :T")
    rank = int(rank)
    kmax = min(rows_cols(F.factors))
    if not 0 <= rank <= kmax:
        raise ArgumentError(f"rank must be between zero and {kmax}")
    if rows_cols(_as_matrix(B))[0] != rank:
        raise DimensionMismatch("solve_r! destination row count must equal rank")
    if rank == 0:
        return B
    d = torch.diagonal(F.factors, dim1=1, dim2=2)[:, :rank]
    zero = torch.all(d == 0, dim=0).cpu().numpy()
    if zero.any():
        raise SingularException(int(np.flatnonzero(zero)[0]) + 1)
    lead = F.factors[:, :rank, :rank]
    if B.dim() == 2:
        trsv_(B, lead, uplo="upper", trans=trans, diag="nonunit", config=config)
    else:
        trsm_(B, lead, side="left", uplo="upper", trans=trans, diag="nonunit", config=config)
    return B


def cholesky_(A, *, check=True, config=None):
    """cholesky! (src/factorizations/cholesky.jl:68-107)."""
    N = _limbs(A)
    info = ctypes.c_int64(0)
    _raise(L.load().mfla_potrf(_handle(), N, _mat(A), ctypes.byref(info), _cfg(config)))
    info = int(info.value)
    if check and info == -1:
        raise DomainError("cholesky!: input matrix contains non-finite entries")
    if check and info > 0:
        raise PosDefException(info)
    return MFCholesky(A, info)


def lu_(A, *, check=True, config=None, workspace=None):
    """lu! (src/factorizations/lu.jl:128-195)."""
    N = _limbs(A)
    m, n = rows_cols(A)
    ipiv = torch.empty(min(m, n), dtype=torch.int64, device=A.device)
    om = np.zeros(4)
    info = ctypes.c_int64(0)
    ws = workspace._p if workspace is not None else None
    _raise(L.load().mfla_getrf(_handle(), N, _mat(A), ctypes.c_void_p(ipiv.data_ptr()), _sp(om), ctypes.byref(info), _cfg(config), ws))
    info = int(info.value)
    if check and info == -1:
        raise DomainError("lu!: input matrix contains non-finite entries")
    if check and info > 0:
        raise SingularException(info)
    return MFLU(A, ipiv, info, om[:N].copy())


def ldlt_(A, *, check=True, config=None, workspace=None):
    """ldlt! (src/factorizations/ldlt.jl:654-717)."""
    N = _limbs(A)
    n = rows_cols(A)[0]
    dsub = torch.zeros((N, n), dtype=torch.float64, device=A.device)
    piv = torch.empty(n, dtype=torch.int64, device=A.device)
    blocks = torch.empty(n, dtype=torch.uint8, device=A.device)
    om = np.zeros(4)
    info = ctypes.c_int64(0)
    ws = workspace._p if workspace is not None else None
    _raise(L.load().mfla_sytrf(_handle(), N, _mat(A), _mat(dsub), ctypes.c_void_p(piv.data_ptr()), ctypes.c_void_p(blocks.data_ptr()),
                               _sp(om), ctypes.byref(info), _cfg(config), ws))
    info = int(info.value)
    if check and info == -1:
        raise DomainError("ldlt!: input matrix contains non-finite entries")
    if check and info != 0:
        raise SingularException(info)
    return MFLDLT(A, dsub, piv, blocks, info, om[:N].copy())


def ldlt_noalloc_(A, *, check=True, config=None):
    """ldlt! through mfla_sytrf_noalloc: the same factorization (bit for
    bit) with the library's static per-GPU device scratch instead of any
    allocation during the call (the outputs dsub, pivots and blocks are
    allocated here, by the caller, as for ldlt_)."""
    N = _limbs(A)
    n = rows_cols(A)[0]
    dsub = torch.zeros((N, n), dtype=torch.float64, device=A.device)
    piv = torch.empty(n, dtype=torch.int64, device=A.device)
    blocks = torch.empty(n, dtype=torch.uint8, device=A.device)
    om = np.zeros(4)
    info = ctypes.c_int64(0)
    _raise(L.load().mfla_sytrf_noalloc(_handle(), N, _mat(A), _mat(dsub), ctypes.c_void_p(piv.data_ptr()),
                                       ctypes.c_void_p(blocks.data_ptr()), _sp(om), ctypes.byref(info), _cfg(config)))
    info = int(info.value)
    if check and info == -1:
        raise DomainError("ldlt!: input matrix contains non-finite entries")
    if check and info != 0:
        raise SingularException(info)
    return MFLDLT(A, dsub, piv, blocks, info, om[:N].copy())


# ----------------------------------------------------------------- solves
def _as_matrix(t):
    return t.unsqueeze(1) if t.dim() == 2 else t


def ldiv_(dest, F, source=None, *, config=None):
    """ldiv!(x, F) / ldiv!(x, F, b) (src/solve.jl:23-423). Validates first;
    a failed factor raises before the destination is touched."""
    if isinstance(F, MFCholesky):
        if F.info != 0:
            raise PosDefException(F.info)
    elif F.info != 0:
        raise SingularException(F.info)
    N = _limbs(dest, F.factors)
    if source is not None:
        if source.shape != dest.shape:
            raise DimensionMismatch("source and destination dimensions differ")
        if source.data_ptr() != dest.data_ptr():
            dest.copy_(source)
    X = _as_matrix(dest)
    lib = L.load()
    h = _handle()
    if isinstance(F, MFCholesky):
        _raise(lib.mfla_potrs(h, N, _mat(F.factors), F.info, _mat(X), _cfg(config)))
    elif isinstance(F, MFLU):
        _raise(lib.mfla_getrs(h, N, _mat(F.factors), ctypes.c_void_p(F.ipiv.data_ptr()), F.info, _mat(X), _cfg(config)))
    elif isinstance(F, MFLDLT):
        _raise(lib.mfla_sytrs(h, N, _mat(F.factors), _mat(F.dsub), ctypes.c_void_p(F.pivots.data_ptr()),
                              ctypes.c_void_p(F.blocks.data_ptr()), F.info, _mat(X), _cfg(config)))
    else:
        raise ArgumentError("unsupported factorization")
    return dest


def solve(F, rhs, *, config=None):
    """solve(F, rhs) (src/solve.jl:555-565): ldiv! on a copy."""
    return ldiv_(rhs.clone(), F, config=config)


def refinement_correction_(delta, F, r, *, config=None):
    """refinement_correction! (src/residual.jl:302-332): exactly one ldiv!."""
    return ldiv_(delta, F, r, config=config)


_UPLO_R = {"general": 2, "lower": 0, "upper": 1}


def residual_(r, A, x, b, *, uplo="general", config=None, workspace=None):
    """residual! (src/residual.jl:55-125): r = b - A*x."""
    N = _limbs(r, A, x, b)
    u = _sym(uplo, _UPLO_R, "uplo must be :general, :lower, or :upper")
    lib = L.load()
    if r.dim() == 2:
        _raise(lib.mfla_residual_vector(_handle(), N, u, _mat(r), _mat(A), _mat(x), _mat(b), _cfg(config)))
    else:
        _raise(lib.mfla_residual(_handle(), N, u, _mat(r), _mat(A), _mat(x), _mat(b), _cfg(config)))
    return r


def residual_mixed_(r, A, x, b, *, uplo="general", config=None):
    """residual_mixed! (src/residual.jl:431-522)."""
    NS = _limbs(A, x, b)
    NR = r.shape[0]
    u = _sym(uplo, _UPLO_R, "uplo must be :general, :lower, or :upper")
    _raise(L.load().mfla_residual_mixed(_handle(), NS, NR, u, _mat(_as_matrix(r)), _mat(A), _mat(_as_matrix(x)),
                                        _mat(_as_matrix(b)), _cfg(config)))
    return r


def add_mixed_(X, D):
    """X .+= D with D promoted to X's limb count (refinement update)."""
    NR, NS = X.shape[0], D.shape[0]
    _raise(L.load().mfla_add_mixed(_handle(), NR, NS, _mat(_as_matrix(X)), _mat(_as_matrix(D))))
    return X


def normwise_backward_error(A, x, b, r, *, uplo="general"):
    """normwise_backward_error (src/residual.jl:215-269). Returns the limbs of
    the value (vector RHS) or an (N, nrhs) array (matrix RHS)."""
    N = _limbs(A, x, b, r)
    u = _sym(uplo, _UPLO_R, "uplo must be :general, :lower, or :upper")
    X, B, R = _as_matrix(x), _as_matrix(b), _as_matrix(r)
    nrhs = X.shape[1]
    out = np.zeros(N * max(nrhs, 1))
    _raise(L.load().mfla_backward_error(_handle(), N, u, _mat(A), _mat(X), _mat(B), _mat(R), _sp(out)))
    out = out.reshape(N, max(nrhs, 1))[:, :nrhs]
    return out[:, 0].copy() if x.dim() == 2 else out


# ------------------------------------------------- binary64 building blocks
_TRANS = {"N": 0, "T": 1}


def dgemm_(C, A, B, alpha=1.0, beta=0.0, *, trans_a="N", trans_b="N", tri_a="general", tri_b="general"):
    """Binary64 C = alpha*op(A)*op(B) + beta*C on the FP64 matrix cores, using
    the leading limb plane of each operand. tri_a / tri_b declare op(A) / op(B)
    lower or upper triangular (the other triangle is never read)."""
    ta = _sym(trans_a, _TRANS, "trans must be 'N' or 'T'")
    tb = _sym(trans_b, _TRANS, "trans must be 'N' or 'T'")
    ra = _sym(tri_a, _UPLO_R, "triangle must be :general, :lower, or :upper")
    rb = _sym(tri_b, _UPLO_R, "triangle must be :general, :lower, or :upper")
    _raise(L.load().mfla_dgemm(_handle(), ta, tb, ra, rb, _mat(C), _mat(A), _mat(B), float(alpha), float(beta)))
    return C


def dpotrf_inverse_(A, Inv):
    """Binary64 Cholesky of A's lower triangle in place and the inverse of the
    factor into Inv's lower triangle. Returns the status (0, -1, k > 0)."""
    info = ctypes.c_int64(0)
    _raise(L.load().mfla_dpotrf_inverse(_handle(), _mat(A), _mat(Inv), ctypes.byref(info)))
    return int(info.value)


def dgetrf_(A, ipiv):
    """Binary64 LU with partial pivoting of the square A's leading limb plane
    in place (LAPACK dgetrf semantics); ipiv: int64 CUDA tensor of n entries,
    one-based. Returns the status (0, -1 non-finite input, k > 0 first zero
    pivot)."""
    n = rows_cols(A)[0]
    if ipiv.dtype != torch.int64 or not ipiv.is_cuda or ipiv.numel() < n:
        raise ArgumentError("ipiv must be an int64 CUDA tensor of n entries")
    info = ctypes.c_int64(0)
    _raise(L.load().mfla_dgetrf(_handle(), _mat(A), ctypes.c_void_p(ipiv.data_ptr()), ctypes.byref(info)))
    return int(info.value)


# --------------------------------------------------------------- interop
def from_interleaved(dst, src):
    """Copy a Julia-layout (AoS) Float64xN matrix -- a CUDA float64 tensor of
    shape (cols, rows, N) -- into the planar tensor dst."""
    N = _limbs(dst)
    rows, cols = rows_cols(dst)
    assert src.shape[-1] == N and src.is_contiguous()
    _raise(L.load().mfla_from_interleaved(_handle(), N, _mat(dst), ctypes.c_void_p(src.data_ptr()), rows))
    return dst


def to_interleaved(src):
    N = _limbs(src)
    rows, cols = rows_cols(src)
    out = torch.empty((cols, rows, N), dtype=torch.float64, device=src.device)
    _raise(L.load().mfla_to_interleaved(_handle(), N, ctypes.c_void_p(out.data_ptr()), rows, _mat(src)))
    return out


ARITH = {"add": 0, "sub": 1, "mul": 2, "div": 3, "sqrt": 4, "mac": 5, "mulacc_x3": 6, "inv": 7, "msub": 8, "abs": 9, "max": 10,
         "min": 11, "lt": 12}


def arith(op, x, y=None, z=None):
    """Elementwise Float64xN op on the GPU over (N, count) tensors."""
    N, count = x.shape
    out = torch.empty_like(x)
    p = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
    _raise(L.load().mfla_arith(_handle(), N, ARITH[op], count, p(x), p(y), p(z), p(out)))
    return out


def host_arith(op, x, y=None, z=None):
    """The same networks on the host over (N, count) numpy arrays."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    N, count = x.shape
    out = np.zeros_like(x)
    yy = np.ascontiguousarray(y, dtype=np.float64) if y is not None else None
    zz = np.ascontiguousarray(z, dtype=np.float64) if z is not None else None
    _raise(L.load().mfla_host_arith(N, ARITH[op], count, _sp(x), _sp(yy) if yy is not None else None,
                                    _sp(zz) if zz is not None else None, _sp(out)))
    return out
