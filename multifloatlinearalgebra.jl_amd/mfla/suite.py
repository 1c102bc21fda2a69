"""Measurement suite: every operation of the Float64xN path at the sizes of the
benchmark configurations, on one GPU, with the work model each figure is
priced by.

Compute-bound operations report EFT FP64-TFLOP/s = MACs * F_N / t against the
FP64 vector peak (78.6 TFLOP/s), with F_N the FLOPs of one multiply-accumulate
network (2, 30, 111, 239 for N = 1..4), and the FP64 instruction issue rate
(instructions per MAC: 2, 27, 108, 229; 99 on the x3 fused route) against
39.3 T lane-instructions/s. Level-2 operations also report HBM GB/s against
8 TB/s with compulsory bytes 8N(mn + m + n) per right-hand side.

Inputs are seeded synthetic Float64xN matrices: hi limb N(0, 1), limb l =
limb(l-1) * 2^-53 * U(-1, 1), renormalised by two fast_two_sum sweeps.
SPD matrices are R R^T + n I formed with gemm!; the triangle of trsm! is the
Cholesky factor of such a matrix.
"""
from __future__ import annotations

import json
import time

import numpy as np
import torch

from . import api as M

F_N = {1: 2, 2: 30, 3: 111, 4: 239}
INSTR_N = {1: 2, 2: 27, 3: 108, 4: 229}
F_X3_FUSED, INSTR_X3_FUSED = 102, 99      # mulacc_x3: MFLA's default x3 gemm! route
PEAK_TF = 78.6
PEAK_INSTR = 39.3e12
PEAK_HBM = 8.0e12


def rand_dev(N, rows, cols, seed=0):
    g = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.empty((N, cols, rows), dtype=torch.float64, device="cuda")
    x[0] = torch.randn((cols, rows), generator=g, dtype=torch.float64, device="cuda")
    for l in range(1, N):
        u = torch.rand((cols, rows), generator=g, dtype=torch.float64, device="cuda") * 2 - 1
        x[l] = x[l - 1] * 2.0**-53 * u
    for _ in range(2):
        for l in range(N - 1):
            sm = x[l] + x[l + 1]
            e = x[l + 1] - (sm - x[l])
            x[l], x[l + 1] = sm, e
    return x


def spd_dev(N, n, seed=1):
    R = rand_dev(N, n, n, seed)
    S = torch.zeros_like(R)
    M.gemm_(S, R, R.transpose(1, 2).contiguous())
    S[0].diagonal().add_(float(n))
    return S


def _time(fn, reps=1, warm=1):
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps / 1e3


def _compute(name, N, macs, t, instr=None, flops=None, **extra):
    instr = instr or INSTR_N[N]
    flops = flops or F_N[N]
    r = {"case": name, "limbs": N, "ms": round(t * 1e3, 3), "macs": macs,
         "tflops": round(macs * flops / t / 1e12, 3), "frac": round(macs * flops / t / 1e12 / PEAK_TF, 4),
         "issue_frac": round(macs * instr / t / PEAK_INSTR, 4)}
    r.update(extra)
    return r


def _memory(name, N, bytes_, macs, t, **extra):
    r = {"case": name, "limbs": N, "ms": round(t * 1e3, 3), "gbs": round(bytes_ / t / 1e9, 1),
         "hbm_frac": round(bytes_ / t / PEAK_HBM, 4), "tflops": round(macs * F_N[N] / t / 1e12, 3),
         "issue_frac": round(macs * INSTR_N[N] / t / PEAK_INSTR, 4)}
    r.update(extra)
    return r


def case_gemm(N, m, k, n, reps=2, strategy="auto"):
    A, B = rand_dev(N, m, k, 1), rand_dev(N, k, n, 2)
    C = torch.zeros((N, n, m), dtype=torch.float64, device="cuda")
    cfg = M.KernelConfig(gemm_strategy=strategy)
    t = _time(lambda: M.gemm_(C, A, B, config=cfg), reps)
    if strategy == "ozaki":
        # INT8 matrix-core work: digit-pair products over K per output
        S = {1: 11, 2: 19, 3: 27, 4: 34}[N]
        pairs = (S + 1) * (S + 2) // 2
        ops = 2.0 * pairs * m * n * ((k + 63) // 64 * 64)
        return {"case": f"gemm {m}x{k}x{n} (ozaki)", "limbs": N, "ms": round(t * 1e3, 3), "macs": m * k * n,
                "x_equiv_tflops": round(m * k * n * F_N[N] / t / 1e12, 3), "int8_tops": round(ops / t / 1e12, 1),
                "int8_frac": round(ops / t / 5.0e15, 4)}
    fused = N == 3 and strategy in ("auto", "fused")
    return _compute(f"gemm {m}x{k}x{n}", N, m * k * n, t, INSTR_X3_FUSED if fused else None, F_X3_FUSED if fused else None,
                    matrix_equiv_gflops=round(2.0 * m * k * n / t / 1e9, 2))


def case_syrk(N, n, k, reps=2):
    P = rand_dev(N, k, n, 3)
    C = torch.zeros((N, n, n), dtype=torch.float64, device="cuda")
    t = _time(lambda: M.syrk_(C, P), reps)
    return _compute(f"syrk n={n} k={k}", N, k * n * (n + 1) // 2, t)


def case_trsm(N, n, nrhs, order="blocked", reps=1):
    F = M.cholesky_(spd_dev(N, n, 4), config=M.KernelConfig())
    B0 = rand_dev(N, n, nrhs, 5)
    B = B0.clone()
    cfg = M.KernelConfig(trsm_order=order)

    def run():
        B.copy_(B0)
        M.trsm_(B, F.factors, config=cfg)
    t = _time(run, reps)
    return _compute(f"trsm lower n={n} nrhs={nrhs} ({order})", N, nrhs * n * (n - 1) // 2, t)


def case_factor(kind, N, n):
    """One factorization of a seeded matrix: SPD R R^T + n I (cholesky,
    ldlt), random (lu), or the symmetric indefinite R + R^T (ldlt_indefinite:
    interchanges and 2x2 pivots in most panels)."""
    if kind == "lu":
        A0 = rand_dev(N, n, n, 7)
    elif kind == "ldlt_indefinite":
        R = rand_dev(N, n, n, 9)
        A0 = (R + R.transpose(1, 2)).contiguous()
    else:
        A0 = spd_dev(N, n, 6)
    A = A0.clone()
    # kernel loading on a small matrix of the same routine (for LDLT an SPD
    # one: the unblocked route on a slice of the indefinite matrix is held
    # off the GPU)
    small = (A0 if kind != "ldlt_indefinite" else spd_dev(N, 64, 6))[:, :64, :64].contiguous()
    fn = {"cholesky": M.cholesky_, "lu": M.lu_, "ldlt": M.ldlt_, "ldlt_indefinite": M.ldlt_}[kind]
    fn(small.clone(), check=False)           # kernel loading
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fn(A, check=False)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    macs = {"cholesky": n**3 // 6, "lu": n**3 // 3, "ldlt": n**3 // 6, "ldlt_indefinite": n**3 // 6}[kind]
    return _compute(f"{kind} n={n}", N, macs, t)


def case_potrs(N, n, nrhs, reps=2):
    """ldiv!(cholesky(S), B): the forward (bitwise MFLA order) and the blocked
    backward triangular solves; work = 2 nrhs n(n-1)/2 MACs."""
    F = M.cholesky_(spd_dev(N, n, 6))
    B0 = rand_dev(N, n, nrhs, 12)
    B = B0.clone()
    M.ldiv_(B, F)

    def run():
        B.copy_(B0)
        M.ldiv_(B, F)
    t = _time(run, reps)
    return _compute(f"potrs n={n} nrhs={nrhs}", N, nrhs * n * (n - 1), t)


def case_qr(N, m, n, kind="rrqr"):
    """rrqr! / qr! of a seeded m x n Float64xN matrix; work = the Householder
    MACs 2 (m n k - (m + n) k^2 / 2 + k^3 / 3) (k = min(m, n); a projection and an update per reflector and column)."""
    A0 = rand_dev(N, m, n, 40)
    fn = M.rrqr_ if kind == "rrqr" else M.qr_
    fn(rand_dev(N, 64, 32, 41))                      # kernel loading
    A = A0.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fn(A)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    k = min(m, n)
    macs = 2 * (m * n * k - (m + n) * k * k // 2 + k ** 3 // 3)
    return _compute(f"{kind} {m}x{n}", N, macs, t)


def case_gemv(N, n, reps=5):
    A = rand_dev(N, n, n, 8)
    x = rand_dev(N, n, 1, 9).reshape(N, n).contiguous()
    y = torch.zeros((N, n), dtype=torch.float64, device="cuda")
    t = _time(lambda: M.gemv_(y, A, x), reps)
    return _memory(f"gemv n={n}", N, 8 * N * (n * n + 2 * n), n * n, t)


def case_trsv(N, n, reps=2):
    F = M.cholesky_(spd_dev(N, n, 10))
    x0 = rand_dev(N, n, 1, 11).reshape(N, n).contiguous()
    x = x0.clone()

    def run():
        x.copy_(x0)
        M.trsv_(x, F.factors)
    t = _time(run, reps)
    return _memory(f"trsv n={n}", N, 8 * N * (n * (n + 1) // 2 + 2 * n), n * (n - 1) // 2, t)


def case_residual(N, n, nrhs, reps=2):
    A = rand_dev(N, n, n, 12)
    X, B = rand_dev(N, n, nrhs, 13), rand_dev(N, n, nrhs, 14)
    R = torch.zeros_like(B)
    t = _time(lambda: M.residual_(R, A, X, B), reps)
    return _memory(f"residual n={n} nrhs={nrhs}", N, 8 * N * (n * n + 3 * n * nrhs), n * n * nrhs, t)


def case_batched(count=10000, nmin=16, nmax=256, N=3, nrhs=4, seed=3):
    """Config 3: uneven SPD / KKT x3 systems in one persistent launch."""
    from .batched import Batch, autotune, synthetic_systems
    systems = synthetic_systems(count, nmin, nmax, N, nrhs, seed)
    b = Batch(systems, N)
    b.solve()                                      # warm-up
    tune = autotune(b, use_cache=False)            # picks shapes, workgroups and order on this batch
    b.restore()
    g = b.graph(tune["workgroups"])
    torch.cuda.synchronize()
    b.restore()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.replay()
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    info = [r["info"] for r in b.results()]
    macs = 0
    for kind, A, B in systems:
        n = A.shape[1]
        macs += n**3 // 6 + 2 * n * n * B.shape[1] // 2
    r = _compute(f"batched {count} x{N} systems n~U{{{nmin}..{nmax}}}", N, macs, t)
    r["systems_per_s"] = round(count / t, 1)
    r["failed"] = int(sum(1 for i in info if i != 0))
    r["autotuned"] = {"shapes": list(tune["shapes"]), "workgroups": tune["workgroups"], "order": tune["order"],
                      "single_launch_ms": min(x["ms"] for x in tune["trials"] if x["shapes"] == [256])}
    return r


def case_autodiff(batch=32, n=1024, nrhs=32, N=4):
    """Config 4: x = S^{-1} B and logdet(S), forward and pullbacks, over a
    batch of 32 problems (after one warm-up problem); the Cholesky rrule's
    pullback is timed separately over the same factors.
    MACs per problem: cholesky! n^3/6, solve n^2 nrhs, solve pullback 2 n^2
    nrhs (S^{-1} Xbar and Bbar X^T), logdet pullback n^3 (S^{-1} from two
    triangular solves with n right-hand sides); Cholesky pullback 2 n^3
    (L^T Lbar as a full product, then two triangular solves)."""
    from . import autodiff as AD
    Ss = [spd_dev(N, n, 100 + i) for i in range(batch)]
    Bs = [rand_dev(N, n, nrhs, 200 + i) for i in range(batch)]
    W = rand_dev(N, n, nrhs, 300)
    Lbar = rand_dev(N, n, n, 301)
    one = np.zeros(N)
    one[0] = 1.0
    factors = []

    def problem(S, B):
        F = M.cholesky_(S.clone())
        X = M.solve(F, B)
        AD.logdet(F)
        AD.solve_adjoint(F, X, W)
        AD.logdet_adjoint(F, one)
        return F

    problem(Ss[0], Bs[0])                      # warm-up: code-object loading

    def run():
        factors.clear()
        for S, B in zip(Ss, Bs):
            factors.append(problem(S, B))
    t = _time(run, 1, warm=0)
    macs = batch * (n**3 // 6 + n * n * nrhs + 2 * n * n * nrhs + n**3)
    r = _compute(f"autodiff batch={batch} n={n} nrhs={nrhs} (cholesky!, solve, logdet + both pullbacks)", N, macs, t,
                 ms_per_problem=round(t * 1e3 / batch, 2))
    AD.cholesky_adjoint(factors[0], Lbar)      # warm-up
    tc = _time(lambda: [AD.cholesky_adjoint(F, Lbar) for F in factors], 1, warm=0)
    r["cholesky_rrule"] = {"ms_per_problem": round(tc * 1e3 / batch, 2),
                           "frac": round(batch * 2 * n**3 * F_N[N] / tc / 1e12 / PEAK_TF, 4)}
    return r


def run(which=("blas3", "factor", "blas2", "refine", "batched", "autodiff"), out=None):
    rows = []

    def go(fn, *args, **kw):
        # the case is named before it runs (a failure then names it)
        print(f"# starting {fn.__name__}{args}{kw if kw else ''}", flush=True)
        emit(fn(*args, **kw))

    def emit(r):
        rows.append(r)
        line = json.dumps(r)
        print(line, flush=True)
        if out:
            with open(out, "a") as f:
                f.write(line + "\n")

    if "blas3" in which:
        for N in (2, 3, 4):
            go(case_gemm, N, 2048, 2048, 2048)
            go(case_gemm, N, 4096, 4096, 4096, reps=1)
            go(case_gemm, N, 8192, 512, 8192, reps=1)
            go(case_syrk, N, 4096, 1024, reps=1)
            go(case_trsm, N, 4096, 256)
    if "factor" in which:
        for kind in ("cholesky", "lu", "ldlt", "ldlt_indefinite"):
            go(case_factor, kind, 4, 4096 if kind in ("cholesky", "lu") else 2048)
        go(case_potrs, 4, 4096, 64)
    if "blas2" in which:
        for N in (2, 3, 4):
            go(case_gemv, N, 4096)
            go(case_trsv, N, 4096)
            go(case_residual, N, 4096, 64)
    if "refine" in which:
        from .refine import bench_refine
        for kind in ("cholesky", "lu"):
            r = bench_refine(kind=kind)
            r["case"] = f"refine {kind} n=4096 nrhs=64"
            emit(r)
    if "ozaki" in which:
        for N in (2, 3, 4):
            go(case_gemm, N, 4096, 4096, 4096, reps=2, strategy="ozaki")
        from .refine import bench_refine
        for kind in ("cholesky", "lu"):
            r = bench_refine(kind=kind, residual="ozaki")
            r["case"] = f"refine {kind} n=4096 nrhs=64 (ozaki residual)"
            emit(r)
    if "qr" in which:
        for N in (2, 4):
            go(case_qr, N, 500, 40)
            go(case_qr, N, 1024, 256)
        go(case_qr, 4, 1024, 1024)
        go(case_qr, 4, 1024, 1024, kind="qr")
    if "batched" in which:
        go(case_batched)
    if "autodiff" in which:
        go(case_autodiff)
    return rows
