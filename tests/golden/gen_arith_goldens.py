"""Writes tests/golden/arith_exact.json: seeded Float64xN operands (limb bit
patterns) for add/mul/div/sqrt and the exact value of each operation as a
512-bit mpmath decimal string. The oracle and the library are checked against
these exact values (error bounds), not against bits: no MultiFloats.jl output
exists for x2/x4 in the reference tree.

Also writes tests/golden/mulacc_x3_proof_vectors.json, the eight fixed limb
vectors of the reference's test/mulacc_x3_proof_vectors.jl:1-58 (data only).

Run: python tests/golden/gen_arith_goldens.py
"""
import json
import os
import sys

import mpmath
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from mfutil import rand_mf  # noqa: E402

mpmath.mp.prec = 512

PROOF_VECTORS = [
    ("ordinary", (0x3FF0000000000000, 0x3C90000000000000, 0x3930000000000000), (0x3FE8000000000000, 0xBC80000000000000, 0x3910000000000000),
     (0xBFF4000000000000, 0x3C80000000000000, 0xB910000000000000), (0x3FB0000000000008, 0x3910000000000000, 0x0000000000000000)),
    ("exact_cancel", (0xBFF0000000000000, 0xBC90000000000000, 0xB930000000000000), (0x3FF0000000000000, 0, 0),
     (0x3FF0000000000000, 0x3C90000000000000, 0x3930000000000000), (0, 0, 0)),
    ("wide_cancel", (0xC310000000000000, 0x3F50000000000000, 0x3BE0000000000000), (0x58F0000000000000, 0x5530000000000000, 0x5170000000000000),
     (0x2A10000000000000, 0xA650000000000000, 0x2290000000000000), (0x3F50000000000000, 0x3BE0800000000000, 0)),
    ("near_underflow", (0x8170000000000000, 0x0000000000004000, 0), (0x20B0000000000000, 0x1CF0000000000000, 0x1930000000000000),
     (0x20B0000000000000, 0x9CF0000000000000, 0x1930000000000000), (0x0000000000004000, 0, 0)),
    ("near_overflow", (0xFE70000000000000, 0x7AB0000000000000, 0x76F0000000000000), (0x5F30000000000000, 0x5B70000000000000, 0x57B0000000000000),
     (0x5F30000000000000, 0xDB70000000000000, 0x57B0000000000000), (0x7AB0000000000000, 0x7700000000000000, 0)),
    ("alternating", (0xC008000000000000, 0x3CAFFFFFFFFFFFFF, 0), (0xBFF8000000000000, 0x3C90000000000000, 0x3920000000000000),
     (0x4000000000000000, 0xBCA0000000000000, 0x3940000000000000), (0xC017FFFFFFFFFFFF, 0xBCBC000000000001, 0x3950000000000000)),
    ("tiny_accumulator", (0x0030000000000004, 0, 0), (0x3370000000000000, 0x2FB0000000000000, 0x2BF0000000000000),
     (0x8DF0000000000000, 0x0A30000000000000, 0x8670000000000000), (0x816FFFFE00000000, 0x0000000000000010, 0)),
    ("zero_product", (0, 0, 0), (0, 0, 0), (0x4090000000000000, 0, 0), (0, 0, 0)),
]


def exact(x, i):
    return mpmath.fsum(mpmath.mpf(float(v)) for v in x[:, i])


def main():
    rng = np.random.default_rng(20260812)
    out = {"prec": 512, "cases": []}
    for N in (2, 3, 4):
        cnt = 24
        a = rand_mf(rng, N, cnt, spread=30)
        b = rand_mf(rng, N, cnt, spread=30)
        # near-cancellation operands for add
        b[:, :6] = -a[:, :6]
        b[N - 1, :6] += a[N - 1, :6] * 2.0**-7
        pos = a * np.sign(a[0])
        for op in ("add", "mul", "div", "sqrt"):
            x = pos if op == "sqrt" else a
            for i in range(cnt):
                X, Y = exact(x, i), exact(b, i)
                val = {"add": X + Y, "mul": X * Y, "div": X / Y, "sqrt": mpmath.sqrt(X)}[op]
                out["cases"].append({
                    "N": N, "op": op,
                    "x": [int(v) for v in x[:, i].view(np.uint64)],
                    "y": [int(v) for v in b[:, i].view(np.uint64)],
                    "exact": mpmath.nstr(val, 160, strip_zeros=False),
                    "scale": mpmath.nstr(abs(X) + abs(Y) if op == "add" else abs(val), 20),
                })
    with open(os.path.join(HERE, "arith_exact.json"), "w") as f:
        json.dump(out, f, indent=0)
    pv = [{"name": n, "acc": [int(v) for v in acc], "x": [int(v) for v in x], "y": [int(v) for v in y], "output": [int(v) for v in o]}
          for (n, acc, x, y, o) in PROOF_VECTORS]
    with open(os.path.join(HERE, "mulacc_x3_proof_vectors.json"), "w") as f:
        json.dump(pv, f, indent=1)


if __name__ == "__main__":
    main()
