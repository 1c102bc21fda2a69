"""GEMM tile calibration: the device counterpart of MFLA's route profiles
(src/calibration.jl:16-335: machine_fingerprint, default_gemm_profile,
with_gemm_profile, profile_compatible, calibrate_gemm).

MFLA calibrates its CPU direct-vs-packed route and panel geometry; here the
tunable is the gemm! output tile of the device kernel (KernelConfig.gemm_tile:
0 = 16x16, 1 = 32x32, 2 = 32x64, 3 = 64x64 for one or two limbs). Every tile
computes each element with the same ascending-k chain, so calibration is a
pure speed choice: like calibrate_gemm, a candidate is accepted only if its
output is bitwise equal to the automatic choice's and it beats it by
`minimum_speedup` at the largest tested size and every larger tested size
from its crossover on (the suffix rule of calibration.jl:178-335). Nothing is
installed globally; a profile is applied with with_gemm_profile.
"""
from __future__ import annotations

import json
import time
from dataclasses import asdict, dataclass, field, replace
from typing import Optional

import numpy as np
import torch

from . import _lib as L
from .api import ArgumentError, KernelConfig, gemm_

TILES = (0, 1, 2, 3)


def machine_fingerprint() -> dict:
    """Stable identity of the device and library (calibration.jl:16-28)."""
    if not torch.cuda.is_available():
        raise ArgumentError("machine_fingerprint needs the GPU the profile is for")
    p = torch.cuda.get_device_properties(0)
    return {
        "device": p.name,
        "arch": getattr(p, "gcnArchName", ""),
        "compute_units": int(p.multi_processor_count),
        "library_version": int(L.load().mfla_version()),
        "hip": str(torch.version.hip),
    }


@dataclass
class GemmProfile:
    """calibration.jl GemmProfile: the resolved gemm! tile for one limb count.
    tile = -1 keeps the automatic choice; crossover = smallest calibrated
    size from which the tile is used (larger outputs keep it)."""
    limbs: int
    tile: int = -1
    crossover: int = 0
    source: str = "builtin"
    fingerprint: dict = field(default_factory=dict)
    measurements: list = field(default_factory=list)

    def to_json(self) -> str:
        return json.dumps(asdict(self))

    @staticmethod
    def from_json(text: str) -> "GemmProfile":
        d = json.loads(text)
        return GemmProfile(**d)


def default_gemm_profile(limbs: int) -> GemmProfile:
    """calibration.jl:39-54: the built-in profile (automatic tile choice)."""
    if limbs not in (1, 2, 3, 4):
        raise ArgumentError("only 1-4 limb MultiFloats are supported")
    fp = machine_fingerprint() if torch.cuda.is_available() else {}
    return GemmProfile(limbs=limbs, fingerprint=fp)


def profile_compatible(profile: GemmProfile, current: Optional[dict] = None) -> bool:
    """calibration.jl:100-115: every identity field of the recorded
    fingerprint matches the current device and library."""
    cur = current if current is not None else machine_fingerprint()
    keys = ("device", "arch", "compute_units", "library_version")
    return all(profile.fingerprint.get(k) == cur.get(k) for k in keys)


def with_gemm_profile(config: Optional[KernelConfig], profile: GemmProfile, *, strict=True,
                      current: Optional[dict] = None) -> KernelConfig:
    """calibration.jl:63-90: a copy of config with the profile's GEMM tile."""
    if strict and not profile_compatible(profile, current):
        raise ArgumentError("GemmProfile fingerprint does not match the current device; pass strict=False to force application")
    base = config if config is not None else KernelConfig()
    return replace(base, gemm_tile=int(profile.tile))


def _rand(N, rows, cols, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.empty((N, cols, rows), dtype=torch.float64, device="cuda")
    x[0] = torch.randn((cols, rows), generator=g, dtype=torch.float64, device="cuda")
    for l in range(1, N):
        x[l] = x[l - 1] * 2.0**-53 * (torch.rand((cols, rows), generator=g, dtype=torch.float64, device="cuda") * 2 - 1)
    return x


def _median_ms(fn, samples):
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(max(samples, 1)):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append((time.perf_counter() - t0) * 1e3)
    return float(np.median(ts))


def calibrate_gemm(limbs: int, sizes=(1024, 2048), samples=3, candidates=TILES, minimum_speedup=1.05):
    """calibration.jl:178-335 for the device tile. Times gemm! at each square
    size for the automatic choice and every candidate tile, rejects any
    candidate whose output differs from the automatic one in any bit, and
    returns (profile, measurements)."""
    if limbs not in (1, 2, 3, 4):
        raise ArgumentError("only 1-4 limb MultiFloats are supported")
    sizes = sorted(int(s) for s in sizes)
    if not sizes or sizes[0] < 1:
        raise ArgumentError("sizes must be positive")
    meas = []
    ok = {t: True for t in candidates}
    speed = {t: {} for t in candidates}
    for n in sizes:
        A, B = _rand(limbs, n, n, 11), _rand(limbs, n, n, 12)
        C = torch.zeros((limbs, n, n), dtype=torch.float64, device="cuda")
        auto_ms = _median_ms(lambda: gemm_(C, A, B), samples)
        ref = C.clone()
        meas.append({"size": n, "tile": -1, "ms": round(auto_ms, 4)})
        for t in candidates:
            cfg = KernelConfig(gemm_tile=t)
            ms = _median_ms(lambda: gemm_(C, A, B, config=cfg), samples)
            same = bool(torch.equal(C.view(torch.int64), ref.view(torch.int64)))
            ok[t] = ok[t] and same
            speed[t][n] = auto_ms / ms
            meas.append({"size": n, "tile": t, "ms": round(ms, 4), "bitwise_equal": same})
    largest = sizes[-1]
    best, best_speed = -1, minimum_speedup
    for t in candidates:
        if ok[t] and speed[t][largest] >= best_speed:
            best, best_speed = t, speed[t][largest]
    crossover = 0
    if best >= 0:
        crossover = largest
        for n in reversed(sizes):          # suffix rule: every size from the crossover on clears the margin
            if speed[best][n] >= minimum_speedup:
                crossover = n
            else:
                break
    prof = GemmProfile(limbs=limbs, tile=best, crossover=crossover, source="calibrated", fingerprint=machine_fingerprint(),
                       measurements=meas)
    return prof, meas
