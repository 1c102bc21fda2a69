"""Benchmark of the Float64xN dense path on one MI355X.

Headline (the JSON `value`): Float64x4 gemm! C = A*B at n = 4096, inputs
resident in HBM, reported in FP64-TFLOP/s with EFT FLOPs counted:
    FLOPs = m*k*n * F_4, F_4 = 239 (x4 add network 124 adds + x4 product
    network 89 adds + 6 muls + 10 FMAs counted as 2; 229 FP64 instructions).
One step = one gemm! call = one kernel launch. The roofline entry prices the
kernel against MI355X's FP64 peak (78.6 TFLOP/s); `issue_frac` gives FP64
instruction issue against 39.3 T lane-instructions/s.

The run also reports the solve comparison (skip with --no-refine): FP64 Cholesky + Float64x4 iterative
refinement against the plain Float64x4 Cholesky solve on the same SPD matrix
(A = Q diag(sigma) Q^T, kappa = 1e8), reporting both times and backward errors.

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
MEASURED_TRAFFIC = {(4, 4096): {"bytes": 39.29e9, "fetch_kib": 18923466, "write_kib": 524288, "fetch_scale": 2.0,
                                "write_scale": 1.0}}


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


def cpu_baseline(A, B, N, n, ncols, threads):
    """The oracle (C++ restatement of MFLA's CPU GEMM, OpenMP over output
    columns) on a column panel of the same product."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.lib()
    Bs = np.ascontiguousarray(B[:, :ncols, :])
    C = np.zeros((N, ncols, n))
    t0 = time.perf_counter()
    O.gemm(C, A, Bs, 1.0, 0.0, fused=True, threads=threads)
    dt = time.perf_counter() - t0
    macs = n * n * ncols
    return {
        "value": macs * NETWORK_COST[N][0] / dt / 1e12,
        "unit": "FP64-TFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{ncols} of {n} output columns of the same x{N} gemm (n={n}), {macs:.3g} MACs in {dt:.2f} s",
    }


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
    flops, instr = NETWORK_COST[N]
    tflops = macs * flops / (ms * 1e-3) / 1e12
    issue = macs * instr / (ms * 1e-3) / PEAK_FP64_INSTR
    return {"A": A, "B": B, "ms": ms, "wall": wall, "tflops": tflops, "issue_frac": issue, "macs": macs}


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
            "x_equiv_tflops": round(float(n) ** 3 * NETWORK_COST[N][0] / (ms * 1e-3) / 1e12, 2),
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
    flops, instr = NETWORK_COST[args.limbs]
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
        "roofline": {"bound": "mfma", "achieved": round(g["tflops"], 4), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(g["tflops"] / PEAK_FP64_TFLOPS, 4),
                     "traffic": MEASURED_TRAFFIC.get((args.limbs, args.n), {}).get("bytes"),
                     "compute_unit": "FP64 VALU (MI355X FP64 vector peak = FP64 matrix peak = 78.6 TFLOP/s)",
                     "issue_frac": round(g["issue_frac"], 4),
                     "compulsory_bytes": 8 * args.limbs * 3 * args.n * args.n},
    }
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        out["cpu_baseline"] = cpu_baseline(g["A"], g["B"], args.limbs, args.n, args.cpu_cols, threads)
    if not args.no_refine:
        from mfla import refine as R

        out["refine"] = R.bench_refine(n=args.n)
    if not args.no_ozaki:
        out["ozaki_gemm"] = bench_ozaki(args, M, torch, g)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
