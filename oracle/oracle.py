"""CPU oracle for the MFLA hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module, and only as the checker (or as the timed CPU
baseline). The product library never imports it.

It binds ``oracle/liboracle.so`` (built from ``mfla_oracle.cpp`` and
``exact.cpp`` by ``oracle/Makefile``), a C++ restatement of the CPU algorithms
of MultiFloatLinearAlgebra.jl with MFLA's per-element operation order.

Array convention (shared with the product): a Float64xN matrix of shape (m, n)
is a float64 numpy array of shape (N, n, m), C-contiguous -- limb planes, each
column-major -- so element (i, j), limb l is ``a[l, j, i]``. Vectors have shape
(N, m). Pivot vectors and status codes are one-based, as in MFLA.

Parity pinning per function (what checks the restatement itself):

* ``arith`` two_sum/fast_two_sum/two_prod -- exactness invariants vs mpmath.
* x3 add / mul / fused mulacc -- the 8 committed proof vectors
  (test/mulacc_x3_proof_vectors.jl:1-58) plus mpmath accuracy.
* x4 add -- topology of benchmark/ordered_twosum.jl:46-81; mpmath accuracy and
  normalisation fixtures (tests/golden/).
* x2 add/mul, x4 mul, div, inv, sqrt -- PARITY UNPINNED against MultiFloats.jl
  bits (the library is not in the reference tree); pinned by mpmath accuracy,
  exactness on representable cases and closed-form identities only.
* geqrf / geqp3 / ormqr (qr!, rrqr!, apply_q!) -- pinned by the reference's
  rank / permutation known-answer tests (test/runtests.jl:753-930,
  test/qr_delayed_norm.jl), by LAPACK's float64 dgeqrf / dgeqp3 (same
  reflector convention: factors, tau and pivot order agree to float64
  accuracy on well-separated norms) and by exact orthogonality / QR = AP
  residuals; blocked and unblocked routes agree on ranks.
* exact_axpb / exact_sum_diff (the stand-in for BigFloat-512) -- against
  fractions.Fraction sums of the same limb products in
  tests/test_oracle_exact.py (exponents over 2^-400..2^400 and the binary64
  range edges, full cancellation C = fl(AB), both signs, atrans).
* gemm / syrk / gemmt / trsm (all 16 variants) -- exact results on integer
  data (index maps, alpha / beta, beta == 0 not reading C, untouched upper
  triangle, unit diagonal) and exact-accumulator bounds on multi-limb data
  (tests/test_oracle_blas.py).
* potrf / getrf / sytrf / solves / gemv / symv / residuals / backward error --
  pinned by the reference's known-answer tests (test/adversarial.jl,
  test/runtests.jl) re-created in tests/, by the exact long accumulator and
  by closed forms (exact solves of diagonal systems, Wilkinson growth
  2^(n-1)...).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

c_i64 = ctypes.c_int64
c_int = ctypes.c_int
c_dp = ctypes.POINTER(ctypes.c_double)
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_u8p = ctypes.POINTER(ctypes.c_uint8)


def build():
    """Compile the oracle with its Makefile (g++, -ffp-contract=off)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.orc_potrf.restype = c_i64
        _lib.orc_getrf.restype = c_i64
        _lib.orc_sytrf.restype = c_i64
        _lib.orc_sytrs.restype = c_i64
        _lib.orc_geqp3.restype = c_i64
        _lib.orc_exact_sum_diff.restype = ctypes.c_double
    return _lib


def _ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(c_dp)


def _mat(a):
    """(ptr, ld, ls) of a planar (N, n, m) array."""
    assert a.ndim == 3
    return _ptr(a), c_i64(a.shape[2]), c_i64(a.shape[1] * a.shape[2])


def _vec(a):
    assert a.ndim == 2
    return _ptr(a), c_i64(a.shape[1])


def _scal(v, N):
    v = np.asarray(v, dtype=np.float64).reshape(-1)
    out = np.zeros(N)
    out[: len(v)] = v[:N]
    return out


def _check(rc):
    if rc not in (0,):
        raise RuntimeError(f"oracle call failed: {rc}")


ARITH_OPS = {"add": 0, "sub": 1, "mul": 2, "div": 3, "sqrt": 4, "mac": 5, "mulacc_x3": 6, "inv": 7, "msub": 8}


def arith(op, x, y=None, z=None):
    """Elementwise multi-double op over planar (N, count) arrays."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    N, count = x.shape
    out = np.zeros_like(x)
    yp = _ptr(np.ascontiguousarray(y, dtype=np.float64)) if y is not None else None
    zp = _ptr(np.ascontiguousarray(z, dtype=np.float64)) if z is not None else None
    _check(lib().orc_arith(c_int(N), c_int(ARITH_OPS[op]), c_i64(count), _ptr(x), yp, zp, _ptr(out)))
    return out


def two_sum(a, b, fast=False):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    s, e = np.zeros_like(a), np.zeros_like(a)
    lib().orc_two_sum(c_i64(a.size), _ptr(a), _ptr(b), _ptr(s), _ptr(e), c_int(1 if fast else 0))
    return s, e


def two_prod(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    p, e = np.zeros_like(a), np.zeros_like(a)
    lib().orc_two_prod(c_i64(a.size), _ptr(a), _ptr(b), _ptr(p), _ptr(e))
    return p, e


def gemm(C, A, B, alpha=1.0, beta=0.0, fused=True, threads=0):
    """C <- alpha*A*B + beta*C (gemm.jl:747-774); fused selects the x3 mulacc route."""
    N, k, m = A.shape
    n = B.shape[1]
    assert B.shape == (N, n, k) and C.shape == (N, n, m)
    _check(lib().orc_gemm(c_int(N), c_i64(m), c_i64(n), c_i64(k), *_mat(A), *_mat(B), *_mat(C),
                          _ptr(_scal(alpha, N)), _ptr(_scal(beta, N)), c_int(1 if fused else 0), c_int(threads)))
    return C


def lower_update(C, L, R, alpha=1.0, beta=0.0, ltrans=False, rtrans=False, threads=0):
    """Lower-triangle C <- alpha*left*right^T + beta*C with left(r,k), right(c,k)."""
    N = C.shape[0]
    n = C.shape[2]
    k = L.shape[2] if ltrans else L.shape[1]
    _check(lib().orc_lower_update(c_int(N), c_i64(n), c_i64(k), *_mat(L), c_int(int(ltrans)), *_mat(R), c_int(int(rtrans)),
                                  *_mat(C), _ptr(_scal(alpha, N)), _ptr(_scal(beta, N)), c_int(threads)))
    return C


def syrk(C, P, alpha=1.0, beta=0.0, threads=0):
    """syrk.jl:67-108: lower C <- alpha*P^T*P + beta*C, P is k x n."""
    return lower_update(C, P, P, alpha, beta, ltrans=True, rtrans=True, threads=threads)


def gemmt(C, L, R, alpha=1.0, beta=0.0, threads=0):
    """gemmt.jl:69-108: lower C <- alpha*L*R^T + beta*C."""
    return lower_update(C, L, R, alpha, beta, threads=threads)


def trsm(B, A, alpha=1.0, side="left", uplo="lower", trans="N", diag="nonunit", threads=0):
    N, n_b, m_b = B.shape
    _check(lib().orc_trsm(c_int(N), c_int(side == "right"), c_int(uplo == "upper"), c_int(trans == "T"), c_int(diag == "unit"),
                          c_i64(m_b), c_i64(n_b), *_mat(A), *_mat(B), _ptr(_scal(alpha, N)), c_int(threads)))
    return B


def trmm(B, A, alpha=1.0, side="left", uplo="lower", trans="N", diag="nonunit"):
    N, n_b, m_b = B.shape
    _check(lib().orc_trmm(c_int(N), c_int(side == "right"), c_int(uplo == "upper"), c_int(trans == "T"), c_int(diag == "unit"),
                          c_i64(m_b), c_i64(n_b), *_mat(A), *_mat(B), _ptr(_scal(alpha, N))))
    return B


def dot(x, y):
    """mfdot over planar (N, cols, rows) or (N, n) arrays; returns N limbs."""
    x3 = x.reshape(x.shape[0], 1, -1) if x.ndim == 2 else x
    y3 = y.reshape(y.shape[0], 1, -1) if y.ndim == 2 else y
    N, cols, rows = x3.shape
    out = np.zeros(N)
    _check(lib().orc_dot(c_int(N), c_i64(rows), c_i64(cols), *_mat(np.ascontiguousarray(x3)), *_mat(np.ascontiguousarray(y3)),
                         _ptr(out)))
    return out


def syrk_packed(out, P, alpha=1.0, beta=0.0, reduction_first=1, reduction_last=None):
    """out: planar (N, n(n+1)/2) packed lower triangle; P planar (N, n, k)."""
    N, n, k = P.shape
    rl = k if reduction_last is None else reduction_last
    _check(lib().orc_syrk_packed(c_int(N), c_i64(n), c_i64(reduction_first - 1), c_i64(rl - 1), *_mat(P), _ptr(out),
                                 c_i64(out.shape[1]), _ptr(_scal(alpha, N)), _ptr(_scal(beta, N))))
    return out


def trsv(x, A, alpha=1.0, uplo="lower", trans="N", diag="nonunit"):
    N, n = x.shape
    x3 = x.reshape(N, 1, n)
    trsm(x3, A, alpha, "left", uplo, trans, diag, threads=1)
    return x


def gemv(y, A, x, alpha=1.0, beta=0.0, trans="N"):
    N, n, m = A.shape
    _check(lib().orc_gemv(c_int(N), c_int(trans == "T"), c_i64(m), c_i64(n), *_mat(A), *_vec(x), *_vec(y),
                          _ptr(_scal(alpha, N)), _ptr(_scal(beta, N))))
    return y


def symv(y, A, x, alpha=1.0, beta=0.0, uplo="lower"):
    N, n, _ = A.shape
    _check(lib().orc_symv(c_int(N), c_int(uplo == "upper"), c_i64(n), *_mat(A), *_vec(x), *_vec(y),
                          _ptr(_scal(alpha, N)), _ptr(_scal(beta, N))))
    return y


def potrf(A, block=16, threads=0):
    N, n, _ = A.shape
    return int(lib().orc_potrf(c_int(N), c_i64(n), *_mat(A), c_i64(block), c_int(threads)))


def getrf(A, block=16, fused=True, threads=0):
    N, n, m = A.shape
    ipiv = np.zeros(min(m, n), dtype=np.int64)
    om = np.zeros(N)
    info = lib().orc_getrf(c_int(N), c_i64(m), c_i64(n), *_mat(A), c_i64(block), ipiv.ctypes.data_as(c_i64p), _ptr(om),
                           c_int(1 if fused else 0), c_int(threads))
    return int(info), ipiv, om


def geqrf(A):
    """qr! (src/factorizations/qr.jl:763-799) in place; returns tau (N, k)."""
    N, n, m = A.shape
    tau = np.zeros((N, min(m, n)))
    lib().orc_geqrf(c_int(N), c_i64(m), c_i64(n), *_mat(A), _ptr(tau), c_i64(tau.shape[1]))
    return tau


def geqp3(A, route="auto"):
    """rrqr! (src/factorizations/qr.jl:692-743) in place. route: "auto" (MFLA's
    choice), "unblocked" or "blocked". Returns (info, tau (N, k), perm one-based)."""
    N, n, m = A.shape
    tau = np.zeros((N, min(m, n)))
    perm = np.zeros(n, dtype=np.int64)
    r = {"auto": 0, "unblocked": 1, "blocked": 2}[route]
    info = lib().orc_geqp3(c_int(N), c_i64(m), c_i64(n), *_mat(A), _ptr(tau), c_i64(tau.shape[1]), perm.ctypes.data_as(c_i64p), c_int(r))
    return int(info), tau, perm


def ormqr(F, tau, B, trans="N"):
    """apply_q! (src/factorizations/qr.jl:938-956): B <- Q B or Q' B in place."""
    N, _, m = F.shape
    B3 = _rhs3(B)
    lib().orc_ormqr(c_int(N), c_int(1 if trans == "T" else 0), c_i64(m), c_i64(tau.shape[1]), *_mat(F), _ptr(tau), c_i64(tau.shape[1]),
                    c_i64(B3.shape[1]), *_mat(B3))
    return B


def sytrf(A, strategy="auto", block=0, crossover=512, threads=0):
    N, n, _ = A.shape
    dsub = np.zeros((N, n))
    piv = np.zeros(n, dtype=np.int64)
    blocks = np.zeros(n, dtype=np.uint8)
    om = np.zeros(N)
    strat = {"auto": 0, "unblocked": 1, "blocked": 2}[strategy]
    info = lib().orc_sytrf(c_int(N), c_i64(n), *_mat(A), c_int(strat), c_i64(block), c_i64(crossover), *_vec(dsub),
                           piv.ctypes.data_as(c_i64p), blocks.ctypes.data_as(c_u8p), _ptr(om), c_int(threads))
    return int(info), dsub, piv, blocks, om


def _rhs3(B):
    return B if B.ndim == 3 else B.reshape(B.shape[0], 1, B.shape[1])


def potrs(F, B, threads=0):
    B3 = _rhs3(B)
    N, nrhs, n = B3.shape
    _check(lib().orc_potrs(c_int(N), c_i64(n), c_i64(nrhs), *_mat(F), *_mat(B3), c_int(threads)))
    return B


def getrs(F, ipiv, B, threads=0):
    B3 = _rhs3(B)
    N, nrhs, n = B3.shape
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int64)
    _check(lib().orc_getrs(c_int(N), c_i64(n), c_i64(nrhs), *_mat(F), ipiv.ctypes.data_as(c_i64p), *_mat(B3), c_int(threads)))
    return B


def sytrs(F, dsub, piv, blocks, B, threads=0):
    B3 = _rhs3(B)
    N, nrhs, n = B3.shape
    piv = np.ascontiguousarray(piv, dtype=np.int64)
    blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
    rc = lib().orc_sytrs(c_int(N), c_i64(n), c_i64(nrhs), *_mat(F), *_vec(dsub), piv.ctypes.data_as(c_i64p),
                         blocks.ctypes.data_as(c_u8p), *_mat(B3), c_int(threads))
    if rc != 0:
        raise RuntimeError(f"sytrs failed: {rc}")
    return B


def backward_error(A, X, B, R, uplo="general"):
    X3, B3, R3 = _rhs3(X), _rhs3(B), _rhs3(R)
    N, n, m = A.shape
    nrhs = X3.shape[1]
    out = np.zeros((N, nrhs))
    u = {"general": 0, "lower": 1, "upper": 2}[uplo]
    _check(lib().orc_backward_error(c_int(N), c_int(u), c_i64(m), c_i64(n), c_i64(nrhs), *_mat(A), *_mat(X3), *_mat(B3),
                                    *_mat(R3), _ptr(out)))
    return out


def residual_mixed(R, A, X, B, uplo="general", threads=0):
    X3, B3, R3 = _rhs3(X), _rhs3(B), _rhs3(R)
    NS, n, m = A.shape
    NR = R3.shape[0]
    u = {"general": 0, "lower": 1, "upper": 2}[uplo]
    _check(lib().orc_residual_mixed(c_int(NS), c_int(NR), c_int(u), c_i64(m), c_i64(n), c_i64(X3.shape[1]), *_mat(A), *_mat(X3),
                                    *_mat(B3), *_mat(R3), c_int(threads)))
    return R


def exact_axpb(A, B, C, sab=1, sc=1, rows=None, cols=None, atrans=False):
    """Exact sc*C[i,j] + sab*sum_k op(A)[i,k]*B[k,j] over sampled (i, j).

    Returns (value, scale) rounded to binary64, scale = |C| + sum |A||B|.
    """
    N = A.shape[0]
    B3 = _rhs3(B)
    k = B3.shape[2]
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    val = np.zeros(len(rows))
    scale = np.zeros(len(rows))
    if C is not None:
        C3 = _rhs3(C)
        cp, cld, cls = _mat(C3)
    else:
        cp, cld, cls = None, c_i64(0), c_i64(0)
    _check(lib().orc_exact_axpb(c_int(N), c_i64(k), *_mat(A), c_int(int(atrans)), *_mat(B3), cp, cld, cls, c_int(sab), c_int(sc),
                                rows.ctypes.data_as(c_i64p), cols.ctypes.data_as(c_i64p), c_i64(len(rows)), _ptr(val), _ptr(scale)))
    return val, scale


def exact_sum_diff(x, y):
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    y = np.ascontiguousarray(y, dtype=np.float64).reshape(-1)
    return lib().orc_exact_sum_diff(_ptr(x), c_i64(x.size), _ptr(y), c_i64(y.size))


def threads():
    return int(lib().orc_threads())
