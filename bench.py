"""Benchmark of the Float64xN dense path on one MI355X.

Headline (the JSON `value`): Float64x4 gemm! C = A*B at n = 4096, inputs
resident in HBM, reported in FP64-TFLOP/s with EFT FLOPs counted:
    FLOPs = m*k*n * F_4, F_4 = 239 (x4 add network 124 adds + x4 product
    network 89 adds + 6 muls + 10 FMAs counted as 2; 229 FP64 instructions).
The traffic figure is the rocprofv3 FETCH_SIZE + WRITE_SIZE of one dispatch of
this kernel (copy-calibrated), measured separately and carried as a constant.
One step = one gemm! call = one kernel launch. The roofline entry prices the
kernel against MI355X's FP64 peak (78.6 TFLOP/s); `issue_frac` gives FP64
instruction issue against 39.3 T lane-instructions/s.

The run also reports the solve comparison (skip with --no-refine): FP64
Cholesky (LU) + Float64x4 iterative refinement against the plain Float64x4
Cholesky (LU) solve on the same SPD A = Q diag(sigma) Q^T (general A = U
diag(sigma) V^T), Q / U / V Haar-random orthogonal, kappa = 1e8, 64
right-hand sides: both times, the library's backward errors and -- in the
cpu_baseline leg, after all timed regions -- the backward errors with the
residual summed exactly by the oracle.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 4096] [--no-refine] [--no-cpu-baseline]
       python bench.py --suite blas3,factor,blas2,refine,ozaki,qr,batched,autodiff   (one JSON line per case)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "multifloatlinearalgebra.jl_amd"))

# FP64 operations per multiply-accumulate of each network: (FLOPs, instructions)
NETWORK_COST = {1: (2, 2), 2: (30, 27), 3: (111, 108), 4: (239, 229)}
NETWORK_COST_X3_FUSED = (102, 99)
PEAK_FP64_TFLOPS = 78.6
PEAK_FP64_INSTR = 39.3e12
# HBM bytes per launch of the Float64x4 n=4096 GEMM kernel: rocprofv3 FETCH_SIZE
# and WRITE_SIZE of one dispatch, each scaled by the same counter's reading for
# a plain 512 MiB device copy (FETCH_SIZE x 2.000, WRITE_SIZE x 1.000).
MEASURED_TRAFFIC = {(4, 4096): {"bytes": 39.24e9, "fetch_kib": 18899661, "write_kib": 524288, "fetch_scale": 2.0,
                                "write_scale": 1.0,
                                "source": "profiles/gemm_x4_n4096.md: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of one dispatch, "
                                          "copy-calibrated"}}


def rand_mf(rng, N, rows, cols):
    x = np.zeros((N, cols, rows))
    x[0] = rng.standard_normal((cols, rows))
    for l in range(1, N):
        x[l] = x[l - 1] * 2.0**-53 * rng.uniform(-1.0, 1.0, (cols, rows))
    for _ in range(2):
        for l in range(N - 1):
            s = x[l] + x[l + 1]
            e = x[l + 1] - (s - x[l])
            x[l], x[l + 1] = s, e
    return x


def network_cost(N):
    """(EFT FLOPs, FP64 instructions) per MAC of the network the default
    route runs: x3 gemm! takes MFLA's fused mulacc_x3 route (gemm.jl:124-127)."""
    return NETWORK_COST_X3_FUSED if N == 3 else NETWORK_COST[N]


def cpu_baseline(A, B, N, n, ncols, threads, runs=3):
    """The oracle (C++ restatement of MFLA's CPU GEMM, OpenMP over output
    columns, MFLA's job decomposition) on column panels of the same product:
    the median of `runs` timings with all host threads, and the same with one
    thread on a narrower panel."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.lib()
    flops = network_cost(N)[0]

    def timed(cols, th):
        Bs = np.ascontiguousarray(B[:, :cols, :])
        ts = []
        for _ in range(runs):
            C = np.zeros((N, cols, n))
            t0 = time.perf_counter()
            O.gemm(C, A, Bs, 1.0, 0.0, fused=True, threads=th)
            ts.append(time.perf_counter() - t0)
        t = float(np.median(ts))
        return n * n * cols * flops / t / 1e12, t, ts

    v_all, t_all, ts_all = timed(ncols, threads)
    one_cols = max(1, ncols // 8)
    v_one, t_one, _ = timed(one_cols, 1)
    return {
        "value": round(v_all, 5),
        "unit": "FP64-TFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{ncols} of {n} output columns of the same x{N} gemm (n={n}, {n * n * ncols:.3g} MACs), median of {runs} runs "
                  f"({', '.join(f'{t:.2f}' for t in ts_all)} s) on {threads} OpenMP threads",
        "nproc": os.cpu_count(),
        "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
        "matrix_equiv_gflops": round(2.0 * n * n * ncols / t_all / 1e9, 4),
        "one_thread": {"value": round(v_one, 5), "unit": "FP64-TFLOP/s", "seconds": round(t_one, 3),
                       "sample": f"{one_cols} output columns, median of {runs} runs"},
    }


def exact_backward_error(A, X, B):
    """max over the given right-hand sides of |b - A x|_inf / (|A|_inf |x|_inf
    + |b|_inf), the residual summed exactly by the oracle's long accumulator
    (the stand-in for MFLA's BigFloat-512 check)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    N, n, _ = A.shape
    k = X.shape[1]
    val, _ = O.exact_axpb(A, X, B, sab=-1, sc=1, rows=np.repeat(np.arange(n), k), cols=np.tile(np.arange(k), n))
    rn = np.abs(val).reshape(n, k).max(axis=0)
    an = np.abs(A.sum(0)).sum(axis=0).max()
    xn = np.abs(X.sum(0)).max(axis=1)
    bn = np.abs(B.sum(0)).max(axis=1)
    return float((rn / (an * xn + bn)).max())


def bench_gemm(args, M, torch):
    N, n = args.limbs, args.n
    rng = np.random.default_rng(20260812)
    A = rand_mf(rng, N, n, n)
    B = rand_mf(rng, N, n, n)
    Ad = torch.from_numpy(A).cuda()
    Bd = torch.from_numpy(B).cuda()
    Cd = torch.zeros((N, n, n), dtype=torch.float64, device="cuda")
    cfg = M.KernelConfig(gemm_tile=args.tile)
    for _ in range(args.warmup):
        M.gemm_(Cd, Ad, Bd, config=cfg)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        M.gemm_(Cd, Ad, Bd, config=cfg)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / args.steps
    macs = float(n) ** 3
    flops, instr = network_cost(N)
    tflops = macs * flops / (ms * 1e-3) / 1e12
    issue = macs * instr / (ms * 1e-3) / PEAK_FP64_INSTR
    return {"A": A, "B": B, "ms": ms, "wall": wall, "tflops": tflops, "issue_frac": issue, "macs": macs,
            "matrix_equiv_gflops": 2.0 * macs / (ms * 1e-3) / 1e9}


def bench_ozaki(args, M, torch, g):
    """The same product with gemm_strategy=:ozaki (INT8 matrix cores, exact
    digit products recombined in Float64xN; normwise accuracy, not MFLA's
    rounding order). Reported beside the headline, never as `value`."""
    N, n = args.limbs, args.n
    Ad = torch.from_numpy(g["A"]).cuda()
    Bd = torch.from_numpy(g["B"]).cuda()
    Cd = torch.zeros((N, n, n), dtype=torch.float64, device="cuda")
    cfg = M.KernelConfig(gemm_strategy="ozaki")
    M.gemm_(Cd, Ad, Bd, config=cfg)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    steps = max(1, args.steps)
    e0.record()
    for _ in range(steps):
        M.gemm_(Cd, Ad, Bd, config=cfg)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    S = {1: 11, 2: 19, 3: 27, 4: 34}[N]
    ops = 2.0 * (S + 1) * (S + 2) / 2 * n * n * ((n + 127) // 128 * 128)
    return {"ms": round(ms, 3), "speedup_vs_eft_gemm": round(g["ms"] / ms, 2),
            "int8_tops": round(ops / (ms * 1e-3) / 1e12, 1), "int8_frac_of_5pops": round(ops / (ms * 1e-3) / 5.0e15, 4),
            "x_equiv_tflops": round(float(n) ** 3 * network_cost(N)[0] / (ms * 1e-3) / 1e12, 2),
            "note": "digit-pair INT8 MFMA products; not EFT FLOPs, not the headline"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--limbs", type=int, default=4)
    ap.add_argument("--tile", type=int, default=-1, help="GEMM output tile: -1 auto, 0 16x16, 1 32x32, 2 32x64, 3 64x64 (x1/x2)")
    ap.add_argument("--cpu-cols", type=int, default=16)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-refine", action="store_true")
    ap.add_argument("--no-ozaki", action="store_true")
    ap.add_argument("--suite", default="",
                    help="comma list of suite groups (blas3,factor,blas2,refine,ozaki,qr,batched,autodiff): print one JSON line "
                         "per measured case instead of the headline line")
    ap.add_argument("--suite-out", default="", help="also append the suite's JSON lines to this file")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="MiB copied once by a plain device copy before the GEMM (scales HBM counters)")
    args = ap.parse_args()
    traffic = MEASURED_TRAFFIC.get((args.limbs, args.n), {})
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: this library has no multi-GPU path (runs with --gpus > 1 are not supported)"}))
        sys.exit(2)

    import torch

    import mfla as M

    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU (the HIP library has no CPU fallback)")
    M.load()
    if args.suite:
        from mfla import suite

        suite.run(tuple(args.suite.split(",")), out=args.suite_out or None)
        return
    if args.copy_calibration:
        src = torch.ones(args.copy_calibration * 2**20 // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    g = bench_gemm(args, M, torch)
    flops, instr = network_cost(args.limbs)
    out = {
        "metric": "Float64x4 GEMM n=4096 FP64-TFLOP/s; refined vs plain x4 solve time at equal error",
        "value": round(g["tflops"], 4),
        "unit": "FP64-TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(g["ms"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Float64x4 entries, all limbs nonzero)",
        "config": {"workload": f"Float64x{args.limbs} square gemm! n={args.n}, C = A*B (beta = 0)", "n": args.n,
                   "limbs": args.limbs, "eft_flops_per_mac": flops, "fp64_instr_per_mac": instr},
        "roofline": {"bound": "fp64_valu", "achieved": round(g["tflops"], 4), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(g["tflops"] / PEAK_FP64_TFLOPS, 4),
                     "traffic": traffic.get("bytes"),
                     "traffic_source": traffic.get("source"),
                     "compute_unit": "FP64 VALU: the EFT networks issue v_add_f64 / v_mul_f64 / v_fma_f64 (no MFMA); "
                                     "MI355X FP64 vector peak 78.6 TFLOP/s",
                     "issue_frac": round(g["issue_frac"], 4),
                     "matrix_equiv_gflops": round(g["matrix_equiv_gflops"], 2),
                     "compulsory_bytes": 8 * args.limbs * 3 * args.n * args.n},
    }
    checks = {}
    if not args.no_refine:
        from mfla import refine as R

        out["refine"] = {}
        for kind in ("cholesky", "lu"):
            r = R.bench_refine(n=args.n, kind=kind, keep_cols=None if args.no_cpu_baseline else [0, 21, 42, 63])
            checks[kind] = r.pop("_check", None)
            out["refine"][kind] = r
    if not args.no_ozaki:
        out["ozaki_gemm"] = bench_ozaki(args, M, torch, g)
    if not args.no_cpu_baseline:
        # the cpu_baseline leg: the oracle timed on the host, then used as the
        # checker of the refine block (after every timed region)
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS") or 0) or min(os.cpu_count() or 1, 16)
        out["cpu_baseline"] = cpu_baseline(g["A"], g["B"], args.limbs, args.n, args.cpu_cols, threads)
        for kind, c in checks.items():
            if c is not None:
                for name in ("refined", "plain"):
                    out["refine"][kind][name]["exact_backward_error"] = exact_backward_error(c["A"], c[name], c["B"])
                out["refine"][kind]["exact_check"] = (f"residual b - A x evaluated exactly (oracle long accumulator) on all "
                                                      f"{c['A'].shape[2]} rows of right-hand sides {c['cols']}; "
                                                      "|r|_inf / (|A|_inf |x|_inf + |b|_inf)")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
