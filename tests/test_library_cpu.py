"""CPU checks of the product library (no GPU needed).

* libmfla_hip.so loads and exports every symbol include/mfla_hip.h declares.
* The library's multi-double networks, instantiated on the host
  (mfla_host_arith, the same source the GPU kernels compile), are bitwise equal
  to the oracle's independent restatement, and reproduce the x3 proof vectors.
"""
import ctypes
import json
import os
import re

import numpy as np
import pytest

from mfutil import rand_mf

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _declared_symbols():
    with open(os.path.join(ROOT, "include", "mfla_hip.h")) as f:
        src = f.read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mfla_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    import mfla
    from mfla import _lib

    lib = mfla.load()
    declared = _declared_symbols()
    assert len(declared) >= 30
    for name in declared:
        assert hasattr(lib, name), name
        assert name in _lib.PROTOTYPES, f"{name} has no ctypes prototype"
    assert lib.mfla_version() == 1


def test_config_default_matches_kernelconfig():
    from mfla import _lib

    lib = _lib.load()
    c = _lib.MFConfig()
    lib.mfla_config_default(ctypes.byref(c))
    # src/types.jl:13-27 defaults
    assert (c.reduction_tile, c.column_tile, c.cholesky_block, c.lu_block, c.ldlt_block) == (64, 16, 16, 16, 0)
    assert c.ldlt_blocked_crossover == 512 and c.gemm_strategy == 0 and c.ldlt_strategy == 0


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_host_networks_equal_oracle(oracle, N):
    import mfla

    rng = np.random.default_rng(40 + N)
    a = rand_mf(rng, N, 300, spread=30)
    b = rand_mf(rng, N, 300, spread=30)
    c = rand_mf(rng, N, 300, spread=30)
    b[:, :30] = -a[:, :30]
    pos = a * np.sign(a[0])
    for op in ("add", "sub", "mul", "div", "mac", "msub", "inv"):
        got = mfla.host_arith(op, a, b, c)
        ref = oracle.arith(op, a, b, c)
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), op
    assert np.array_equal(mfla.host_arith("sqrt", pos).view(np.uint64), oracle.arith("sqrt", pos).view(np.uint64))
    if N == 3:
        assert np.array_equal(mfla.host_arith("mulacc_x3", a, b, c).view(np.uint64),
                              oracle.arith("mulacc_x3", a, b, c).view(np.uint64))


def test_host_networks_proof_vectors():
    import mfla

    with open(os.path.join(HERE, "golden", "mulacc_x3_proof_vectors.json")) as f:
        pv = json.load(f)
    f64 = lambda k: np.stack([np.array(v[k], dtype=np.uint64).view(np.float64) for v in pv], axis=1)  # noqa: E731
    acc, x, y, out = f64("acc"), f64("x"), f64("y"), f64("output")
    assert np.array_equal(mfla.host_arith("mac", x, y, acc).view(np.uint64), out.view(np.uint64))
    assert np.array_equal(mfla.host_arith("mulacc_x3", x, y, acc).view(np.uint64), out.view(np.uint64))


def test_argument_errors_before_gpu_work():
    import mfla

    with pytest.raises(mfla.ArgumentError):
        mfla.host_arith("mulacc_x3", np.zeros((2, 3)), np.zeros((2, 3)), np.zeros((2, 3)))
    with pytest.raises(mfla.ArgumentError):
        mfla.KernelConfig(gemm_strategy="bogus")._c()
