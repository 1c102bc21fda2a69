"""Shared test helpers: seeded multi-double generators and comparisons.

Matrices are limb-planar numpy arrays of shape (N, cols, rows) (the layout of
the oracle and of the library); vectors are (N, n).
"""
import zlib

import numpy as np
import pytest

# GPU cases held off the GPU: a run of the GPU suite that included them ended
# its box with a GPU fault (memory fault or hung queue) and printed nothing
# that names the faulting kernel; their code paths are kept, but they are not
# run on a GPU until a cause is shown.
HELD_OFF_GPU = pytest.mark.skip(reason="held off the GPU: a suite run that included this case ended its box with an "
                                       "unattributed GPU fault")


def renorm(x):
    """fast_two_sum sweeps from the top limb down, twice (plain IEEE numpy)."""
    N = x.shape[0]
    for _ in range(2):
        for l in range(N - 1):
            s = x[l] + x[l + 1]
            bp = s - x[l]
            e = x[l + 1] - bp
            x[l] = s
            x[l + 1] = e
    return x


def case_seed(*parts):
    """A seed fixed by the test case's parameters (the same in every process,
    unlike hash() of strings)."""
    return zlib.crc32(repr(parts).encode())


def rand_mf(rng, N, rows, cols=None, spread=0, scale=1.0):
    """Seeded Float64xN entries: hi ~ N(0,1) (times 2^U{-spread..spread}),
    limb_k = limb_{k-1} * 2^-53 * U(-1, 1), renormalised."""
    shape = (cols, rows) if cols is not None else (rows,)
    x = np.zeros((N,) + shape)
    x[0] = rng.standard_normal(shape) * scale
    if spread:
        x[0] *= 2.0 ** rng.integers(-spread, spread + 1, shape)
    for l in range(1, N):
        x[l] = x[l - 1] * 2.0**-53 * rng.uniform(-1.0, 1.0, shape)
    return renorm(x)


def from_double(a, N):
    """Promote a float64 array (rows, cols) or (n,) into planar Float64xN."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 2:
        out = np.zeros((N, a.shape[1], a.shape[0]))
        out[0] = a.T
    else:
        out = np.zeros((N, a.shape[0]))
        out[0] = a
    return out


def to_dense(x):
    """Sum of limbs as float64 (rows, cols) (for rough checks only)."""
    s = x.sum(axis=0)
    return s.T if s.ndim == 2 else s


def bits(x):
    return np.ascontiguousarray(x).view(np.uint64)


def assert_bitwise(a, b, what=""):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    ba, bb = bits(a), bits(b)
    if not np.array_equal(ba, bb):
        bad = np.argwhere(ba != bb)
        raise AssertionError(f"{what}: {len(bad)} limbs differ, first at {bad[0].tolist()}: {a[tuple(bad[0])]!r} vs {b[tuple(bad[0])]!r}")


def eps(N):
    """eps(Float64xN) = 2^(-52 N)."""
    return 2.0 ** (-52 * N)
