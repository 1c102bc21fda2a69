"""GEMM tile profiles (the device counterpart of src/calibration.jl:39-115):
JSON round trip, fingerprint compatibility and with_gemm_profile, on the CPU
(no device calls)."""
import pytest

import mfla as M

FP = {"device": "AMD Instinct MI355X", "arch": "gfx950", "compute_units": 256, "library_version": 1, "hip": "7.0"}


def test_profile_round_trip_and_apply():
    p = M.GemmProfile(limbs=2, tile=3, crossover=1024, source="calibrated", fingerprint=dict(FP),
                      measurements=[{"size": 1024, "tile": 3, "ms": 1.0, "bitwise_equal": True}])
    q = M.GemmProfile.from_json(p.to_json())
    assert q == p
    cfg = M.with_gemm_profile(M.KernelConfig(cholesky_block=8), q, current=dict(FP))
    assert cfg.gemm_tile == 3 and cfg.cholesky_block == 8


def test_incompatible_fingerprint_is_rejected_unless_forced():
    p = M.GemmProfile(limbs=4, tile=2, fingerprint=dict(FP))
    other = dict(FP, compute_units=304)
    assert M.profile_compatible(p, current=dict(FP)) and not M.profile_compatible(p, current=other)
    with pytest.raises(M.ArgumentError):
        M.with_gemm_profile(None, p, current=other)
    assert M.with_gemm_profile(None, p, strict=False, current=other).gemm_tile == 2


def test_builtin_profile_keeps_the_automatic_tile():
    p = M.GemmProfile(limbs=3)
    assert p.tile == -1 and p.source == "builtin"
    assert M.with_gemm_profile(None, p, current={}).gemm_tile == -1
