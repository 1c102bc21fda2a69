"""Batched factor + solve of many small uneven systems (one persistent launch).

MFLA has no batched API: a caller loops ``cholesky!``/``ldlt!`` and
``ldiv!`` over its systems. ``Batch`` packs such a loop's systems into device
pools and runs them all with ``mfla_batched_solve``: every system gets MFLA's
own factorization (bitwise equal to the single-system call) and an in-place
solve of its right-hand sides. The launch allocates nothing and never
synchronises, so ``Batch.graph()`` can capture it (restore + solve) in a HIP
graph and replay it.

The autotuner picks the launch's tile shapes -- the workgroup shape (64, 128
or 256 threads) each size class of the batch runs on, the persistent
workgroup count of each class and the queue order -- by timing candidates on
the batch itself; none changes any result (each system is computed by one
workgroup in a fixed order).
"""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np
import torch

from . import _lib as L
from .api import KernelConfig, _cfg, _handle, _mat, _raise

CHOLESKY, LDLT = 0, 1
ENTRY_DTYPE = np.dtype([("n", "<i8"), ("kind", "<i4"), ("nrhs", "<i4"), ("a_offset", "<i8"), ("b_offset", "<i8"),
                        ("meta_offset", "<i8")])
_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "mfla_batched_tuning.json")


SHAPES = (64, 128, 256)   # workgroup shapes (threads; a system of order n needs n <= threads)


class Batch:
    """Systems given as (kind, A, B) with planar numpy arrays A (N, n, n) and
    B (N, nrhs, n); kind is CHOLESKY (lower triangle read) or LDLT.

    shapes: the workgroup shapes the batch is split over. Each system goes to
    the smallest shape with n <= threads (its size class); every class runs
    as its own persistent launch, so small systems do not hold a 256-thread
    workgroup. (256,) runs the whole batch in one launch."""

    def __init__(self, systems, limbs, order=None, shapes=(256,)):
        self.N = limbs
        self.systems = systems
        n_list = [A.shape[1] for _, A, _ in systems]
        self.count = len(systems)
        ent = np.zeros(self.count, dtype=ENTRY_DTYPE)
        a_off = b_off = m_off = 0
        a_parts, b_parts = [], []
        for i, (kind, A, B) in enumerate(systems):
            n, nrhs = A.shape[1], B.shape[1]
            ent[i] = (n, kind, nrhs, a_off, b_off, m_off)
            a_parts.append(np.ascontiguousarray(A).reshape(-1))
            b_parts.append(np.ascontiguousarray(B).reshape(-1))
            a_off += self.N * n * n
            b_off += self.N * n * nrhs
            m_off += self.N * n
        self.entries_host = ent
        dev = "cuda"
        self.A0 = torch.from_numpy(np.concatenate(a_parts) if a_parts else np.zeros(1)).to(dev)
        self.B0 = torch.from_numpy(np.concatenate(b_parts) if b_parts else np.zeros(1)).to(dev)
        self.A = self.A0.clone()
        self.B = self.B0.clone()
        self.dsub = torch.zeros(max(m_off, 1), dtype=torch.float64, device=dev)
        self.piv = torch.zeros(max(m_off, 1), dtype=torch.int64, device=dev)
        self.blk = torch.zeros(max(m_off, 1), dtype=torch.uint8, device=dev)
        self.info = torch.zeros(max(self.count, 1), dtype=torch.int64, device=dev)
        self.origmax = torch.zeros(max(self.count, 1) * self.N, dtype=torch.float64, device=dev)
        self.counter = torch.zeros(len(SHAPES), dtype=torch.int32, device=dev)
        self.max_n = max(n_list) if n_list else 0
        self.set_order(order, shapes)

    def set_order(self, order=None, shapes=None):
        """Queue order within each size class: None keeps the given order,
        "desc" runs the largest systems first (longest-processing-time
        first); shapes as in the constructor (None keeps the current ones)."""
        shapes = tuple(sorted(set(shapes if shapes is not None else self.shapes)))
        if not shapes or any(t not in SHAPES for t in shapes):
            raise ValueError(f"shapes must be drawn from {SHAPES}")
        ent = self.entries_host
        # size class of each system: the smallest shape that holds it (the
        # largest for oversized systems, which the kernel reports as info -2)
        cls = np.searchsorted(np.asarray(shapes), ent["n"], side="left").clip(max=len(shapes) - 1)
        key = -ent["n"] if order == "desc" else np.zeros(len(ent), dtype=np.int64)
        perm = np.lexsort((key, cls)) if len(ent) else np.zeros(0, dtype=np.int64)
        ent = ent[perm]
        cls = cls[perm]
        self.order = order
        self.shapes = shapes
        self.classes = []
        for c, t in enumerate(shapes):
            idx = np.nonzero(cls == c)[0]
            if len(idx):
                self.classes.append((t, int(idx[0]), int(idx[-1]) + 1))
        self._perm = ent
        self.entries = torch.from_numpy(ent.view(np.uint8).copy()).cuda()
        # info/origmax are indexed by queue position: keep the permutation
        self._qpos = {int(o): i for i, o in enumerate(ent["a_offset"])}

    def _flat(self, t, rows, cols):
        return L.MFMatrix(t.data_ptr(), rows, cols, max(rows, 1), max(rows * cols, 1))

    def restore(self):
        """Copies the pristine inputs back with device kernels (graph-safe)."""
        lib = L.load()
        h = _handle()
        for src, dst in ((self.A0, self.A), (self.B0, self.B)):
            nel = src.numel()
            _raise(lib.mfla_copy(h, 1, L.MFMatrix(dst.data_ptr(), nel, 1, nel, nel), L.MFMatrix(src.data_ptr(), nel, 1, nel, nel)))

    def solve(self, workgroups=0, exact_backward=False, config=None):
        """One persistent launch per size class. workgroups: 0 (the library's
        default per shape), an int for every class, or a dict {threads: n}."""
        if self.max_n > L.load().mfla_batched_max_order():
            raise ValueError(f"batched systems need n <= {L.load().mfla_batched_max_order()}")
        lib, h = L.load(), _handle()
        esize = ENTRY_DTYPE.itemsize
        for c, (t, lo, hi) in enumerate(self.classes):
            wg = workgroups.get(t, 0) if isinstance(workgroups, dict) else workgroups
            _raise(lib.mfla_batched_solve(h, self.N, ctypes.c_void_p(self.entries.data_ptr() + lo * esize), hi - lo,
                                          ctypes.c_void_p(self.A.data_ptr()), ctypes.c_void_p(self.B.data_ptr()),
                                          ctypes.c_void_p(self.dsub.data_ptr()), ctypes.c_void_p(self.piv.data_ptr()),
                                          ctypes.c_void_p(self.blk.data_ptr()), ctypes.c_void_p(self.info.data_ptr() + 8 * lo),
                                          ctypes.c_void_p(self.origmax.data_ptr() + 8 * self.N * lo),
                                          ctypes.c_void_p(self.counter.data_ptr() + 4 * c), int(wg), int(t),
                                          1 if exact_backward else 0, _cfg(config)))

    def graph(self, workgroups=0, exact_backward=False, config=None):
        """Captures restore + solve in a HIP graph; returns the graph (replay())."""
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.restore()
            self.solve(workgroups, exact_backward, config)        # warm up outside capture
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.restore()
            self.solve(workgroups, exact_backward, config)
        torch.cuda.synchronize()
        return g

    # --------------------------------------------------------- results
    def results(self):
        """Per system: dict(info, A (factor), X, pivots, blocks, dsub, original_maximum)."""
        torch.cuda.synchronize()
        A = self.A.cpu().numpy()
        B = self.B.cpu().numpy()
        info = self.info.cpu().numpy()
        om = self.origmax.cpu().numpy()
        dsub = self.dsub.cpu().numpy()
        piv = self.piv.cpu().numpy()
        blk = self.blk.cpu().numpy()
        out = []
        for i, e in enumerate(self.entries_host):
            n, nrhs, ao, bo, mo = int(e["n"]), int(e["nrhs"]), int(e["a_offset"]), int(e["b_offset"]), int(e["meta_offset"])
            q = self._qpos[ao]
            r = {"info": int(info[q]), "A": A[ao: ao + self.N * n * n].reshape(self.N, n, n),
                 "X": B[bo: bo + self.N * n * nrhs].reshape(self.N, nrhs, n), "original_maximum": om[q * self.N: (q + 1) * self.N]}
            if e["kind"] == LDLT:
                r["pivots"] = piv[mo: mo + n].copy()
                r["blocks"] = blk[mo: mo + n].copy()
                r["dsub"] = dsub[mo: mo + self.N * n].reshape(self.N, n)
            out.append(r)
        return out


def _sym_low_limbs(hi, N, rng):
    """Symmetric lower limbs hi * 2^-53k * U(-1, 1) (valid Float64xN values)."""
    out = np.zeros((N,) + hi.shape)
    out[0] = hi
    for l in range(1, N):
        u = rng.uniform(-1.0, 1.0, hi.shape)
        u = (u + u.T) / 2
        out[l] = out[l - 1] * 2.0**-53 * u
    return out


def synthetic_systems(count, nmin=16, nmax=256, N=3, nrhs=4, seed=3):
    """The uneven batch of the benchmark configuration: n ~ U{nmin..nmax},
    alternating SPD R R^T + n I (kind CHOLESKY) and KKT [H A^T; A -1e-8 I]
    with H = R R^T + m I (m x m), A ~ N(0, 1) (p x m), p = round(m / 2),
    m + p = n (kind LDLT); right-hand sides ~ N(0, 1). Planar (N, cols, rows)
    arrays, symmetric in every limb."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        if i % 2 == 0:
            R = rng.standard_normal((n, n))
            A = _sym_low_limbs(R @ R.T + n * np.eye(n), N, rng)
            kind = CHOLESKY
        else:
            m = int(round(2 * n / 3))
            p = n - m
            R = rng.standard_normal((m, m))
            K = np.zeros((n, n))
            K[:m, :m] = R @ R.T + m * np.eye(m)
            Ab = rng.standard_normal((p, m))
            K[m:, :m] = Ab
            K[:m, m:] = Ab.T
            K[m:, m:] = -1e-8 * np.eye(p)
            A = _sym_low_limbs(K, N, rng)
            kind = LDLT
        B = np.zeros((N, nrhs, n))
        B[0] = rng.standard_normal((nrhs, n))
        out.append((kind, A, B))
    return out


def _time(fn, reps=3):
    ts = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
    return float(np.median(ts))


def autotune(batch: Batch, shape_sets=((256,), (128, 256), (64, 256), (64, 128, 256)), scales=(0.5, 1, 2),
             orders=(None, "desc"), use_cache=True):
    """The batched launch's tile-shape autotuner: times every candidate --
    the set of workgroup shapes the batch is split over (size classes of
    64 / 128 / 256 threads), the persistent workgroup count of each class
    (scale x its resident capacity, 256 CUs x 16 waves), and the queue order
    -- on the batch itself (restore + solve, median of 3), and keeps the
    fastest. No candidate changes any result (each system is computed by one
    workgroup in a fixed order). The choice is cached per (limbs, count,
    size sum)."""
    key = f"x{batch.N}:{batch.count}:{int(batch.entries_host['n'].sum())}"
    if use_cache and os.path.exists(_CACHE):
        try:
            with open(_CACHE) as f:
                c = json.load(f)
            if key in c and "shapes" in c[key]:
                e = c[key]
                batch.set_order(e["order"], tuple(e["shapes"]))
                return {"shapes": tuple(e["shapes"]), "workgroups": {int(k): v for k, v in e["workgroups"].items()},
                        "order": e["order"], "cached": True}
        except (OSError, ValueError):
            pass
    best = None
    trials = []
    for shapes in shape_sets:
        for order in orders:
            batch.set_order(order, shapes)
            for sc in scales:
                wg = {t: max(64, int(sc * 256 * 4 * (256 // t))) for t in shapes}
                ms = _time(lambda: (batch.restore(), batch.solve(wg)))
                trials.append({"shapes": list(shapes), "workgroups": wg, "order": order, "ms": round(ms, 3)})
                if best is None or ms < best[0]:
                    best = (ms, shapes, wg, order)
    batch.set_order(best[3], best[1])
    res = {"shapes": best[1], "workgroups": best[2], "order": best[3], "ms": round(best[0], 3), "trials": trials}
    try:
        os.makedirs(os.path.dirname(_CACHE), exist_ok=True)
        c = {}
        if os.path.exists(_CACHE):
            with open(_CACHE) as f:
                c = json.load(f)
        c[key] = {"shapes": list(best[1]), "workgroups": {str(k): v for k, v in best[2].items()}, "order": best[3]}
        with open(_CACHE, "w") as f:
            json.dump(c, f)
    except (OSError, ValueError):
        pass
    return res
