"""Backend adapters for tests/kat_cases.py: the oracle (CPU) and the HIP
library (GPU) behind one small interface over planar numpy arrays."""
import numpy as np


class OracleBackend:
    def __init__(self, O):
        self.O = O

    def arith(self, op, a, b=None):
        return self.O.arith(op, a, b)

    def matvec(self, A, x):
        N, n, m = A.shape
        y = np.zeros((N, m))
        return self.O.gemv(y, A, x, 1.0, 0.0)

    def potrf(self, A, block=16):
        A = A.copy()
        return A, self.O.potrf(A, block)

    def getrf(self, A, block=16):
        A = A.copy()
        info, ipiv, om = self.O.getrf(A, block)
        return A, info, ipiv, om

    def sytrf(self, A, strategy="auto", block=0, crossover=512):
        A = A.copy()
        info, dsub, piv, blocks, om = self.O.sytrf(A, strategy, block, crossover)
        return A, info, dsub, piv, blocks, om

    def potrs(self, F, B):
        return self.O.potrs(F, B.copy())

    def getrs(self, F, ipiv, B):
        return self.O.getrs(F, ipiv, B.copy())

    def sytrs(self, F, dsub, piv, blocks, B):
        return self.O.sytrs(F, dsub, piv, blocks, B.copy())

    def residual(self, A, x, b, uplo="general"):
        r = b.copy()
        if uplo == "general":
            return self.O.gemv(r, A, x, -1.0, 1.0)
        return self.O.symv(r, A, x, -1.0, 1.0, uplo)

    def residual_mixed(self, A, x, b, NR):
        r = np.zeros((NR,) + b.shape[1:])
        return self.O.residual_mixed(r, A, x, b)

    def backward_error(self, A, x, b, r, uplo="general"):
        return self.O.backward_error(A, x, b, r, uplo)[:, 0]

    def absmax_upper(self, F):
        N, n, m = F.shape
        best = None
        for j in range(n):
            for i in range(min(j + 1, m)):
                v = F[:, j, i].copy()
                if np.signbit(v[0]):
                    v = -v
                if best is None or _lt(best, v):
                    best = v
        return best

    def inertia(self, F, dsub, blocks):
        N = F.shape[0]
        pos = neg = zer = 0
        k, n = 0, len(blocks)
        one = lambda v: (1, 0, 0) if _gt0(v) else ((0, 1, 0) if _lt0(v) else (0, 0, 1))  # noqa: E731
        while k < n:
            if blocks[k] == 1:
                p, q, z = one(F[:, k, k])
                k += 1
            elif blocks[k] == 2 and k < n - 1:
                p, q, z = self._in2(N, F[:, k, k], dsub[:, k], F[:, k + 1, k + 1])
                k += 2
            else:
                break
            pos, neg, zer = pos + p, neg + q, zer + z
        return (pos, neg, zer)

    def _in2(self, N, d11, d21, d22):
        # diagnostics.jl:118-134 evaluated with the oracle's networks
        a1 = lambda op, x, y=None: self.O.arith(op, x.reshape(N, 1), None if y is None else y.reshape(N, 1))[:, 0]  # noqa: E731
        ab = lambda v: -v if np.signbit(v[0]) else v  # noqa: E731
        mx = lambda x, y: x if _lt(y, x) else y  # noqa: E731
        scale = mx(mx(ab(d11), ab(d21)), ab(d22))
        if np.all(scale == 0):
            return (0, 0, 2)
        a, b, c = a1("div", d11, scale), a1("div", d21, scale), a1("div", d22, scale)
        det = a1("sub", a1("mul", a, c), a1("mul", b, b))
        tr = a1("add", a, c)
        if _lt0(det):
            return (1, 1, 0)
        if _gt0(det):
            return (2, 0, 0) if _gt0(tr) else (0, 2, 0)
        return (1, 0, 1) if _gt0(tr) else ((0, 1, 1) if _lt0(tr) else (0, 0, 2))

    def permutation(self, piv, blocks):
        n = len(blocks)
        perm = list(range(1, n + 1))
        k = 0
        while k < n:
            if blocks[k] == 1:
                p = piv[k] - 1
                perm[k], perm[p] = perm[p], perm[k]
                k += 1
            elif blocks[k] == 2 and k < n - 1:
                p = piv[k] - 1
                perm[k + 1], perm[p] = perm[p], perm[k + 1]
                k += 2
            else:
                break
        return perm


def _lt(a, b):
    for i in range(len(a)):
        if a[i] < b[i]:
            return True
        if a[i] > b[i]:
            return False
    return False


def _gt0(v):
    return _lt(np.zeros_like(v), v)


def _lt0(v):
    return _lt(v, np.zeros_like(v))


class GpuBackend:
    """Drives the HIP library through the mfla Python mirror."""

    def __init__(self):
        import torch

        import mfla

        self.torch, self.M = torch, mfla

    def d(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a)).cuda()

    def h(self, t):
        self.torch.cuda.synchronize()
        return t.cpu().numpy()

    def arith(self, op, a, b=None):
        return self.h(self.M.arith(op, self.d(a), None if b is None else self.d(b)))

    def matvec(self, A, x):
        N, n, m = A.shape
        y = self.torch.zeros((N, m), dtype=self.torch.float64, device="cuda")
        self.M.gemv_(y, self.d(A), self.d(x), 1.0, 0.0)
        return self.h(y)

    def potrf(self, A, block=16):
        Ad = self.d(A)
        F = self.M.cholesky_(Ad, check=False, config=self.M.KernelConfig(cholesky_block=block))
        return self.h(Ad), F.info

    def getrf(self, A, block=16):
        Ad = self.d(A)
        F = self.M.lu_(Ad, check=False, config=self.M.KernelConfig(lu_block=block))
        self._lu = F
        return self.h(Ad), F.info, F.ipiv.cpu().numpy(), F.original_maximum

    def sytrf(self, A, strategy="auto", block=0, crossover=512):
        Ad = self.d(A)
        F = self.M.ldlt_(Ad, check=False,
                         config=self.M.KernelConfig(ldlt_strategy=strategy, ldlt_block=block, ldlt_blocked_crossover=crossover))
        return self.h(Ad), F.info, self.h(F.dsub), F.pivots.cpu().numpy(), F.blocks.cpu().numpy(), F.original_maximum

    def potrs(self, F, B):
        f = self.M.MFCholesky(self.d(F), 0)
        return self.h(self.M.solve(f, self.d(B)))

    def getrs(self, F, ipiv, B):
        f = self.M.MFLU(self.d(F), self.torch.from_numpy(np.ascontiguousarray(ipiv)).cuda(), 0, np.zeros(F.shape[0]))
        return self.h(self.M.solve(f, self.d(B)))

    def sytrs(self, F, dsub, piv, blocks, B):
        t = self.torch
        f = self.M.MFLDLT(self.d(F), self.d(dsub), t.from_numpy(np.ascontiguousarray(piv)).cuda(),
                          t.from_numpy(np.ascontiguousarray(blocks)).cuda(), 0, np.zeros(F.shape[0]))
        return self.h(self.M.solve(f, self.d(B)))

    def residual(self, A, x, b, uplo="general"):
        r = self.torch.zeros_like(self.d(b))
        self.M.residual_(r, self.d(A), self.d(x), self.d(b), uplo=uplo)
        return self.h(r)

    def residual_mixed(self, A, x, b, NR):
        r = self.torch.zeros((NR,) + tuple(b.shape[1:]), dtype=self.torch.float64, device="cuda")
        self.M.residual_mixed_(r, self.d(A), self.d(x), self.d(b))
        return self.h(r)

    def backward_error(self, A, x, b, r, uplo="general"):
        return self.M.normwise_backward_error(self.d(A), self.d(x), self.d(b), self.d(r), uplo=uplo)

    def absmax_upper(self, F):
        from mfla.factor import _absmax

        return _absmax(self.d(F), 2)

    def inertia(self, F, dsub, blocks):
        t = self.torch
        f = self.M.MFLDLT(self.d(F), self.d(dsub), t.zeros(len(blocks), dtype=t.int64, device="cuda"),
                          t.from_numpy(np.ascontiguousarray(blocks)).cuda(), 0, np.zeros(F.shape[0]))
        r = self.M.factor_inertia(f)
        return (r["positive"], r["negative"], r["zero"])

    def permutation(self, piv, blocks):
        t = self.torch
        f = self.M.MFLDLT(None, None, t.from_numpy(np.ascontiguousarray(piv)), t.from_numpy(np.ascontiguousarray(blocks)), 0, None)
        return list(self.M.factor_permutation(f))
