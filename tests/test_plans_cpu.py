"""gemm_plan / GemmPlan and ldlt_plan / LDLTPlan through the C ABI (host-only
entry points: no GPU needed). The MFLA-side fields restate the reference's
own expectations (test/runtests.jl:307-312, 452-480, 1130-1139, 2516-2518,
2767-2775); the device fields must describe the launch gemm! makes."""
import pytest

import mfla as M


@pytest.mark.parametrize("N", [2, 3, 4])
def test_direct_and_packed_plans(N):
    # runtests.jl:275-312
    n = 23
    d = M.gemm_plan(N, n, n, n, M.KernelConfig(thread_count=2, gemm_strategy="direct", gemm_panel_columns=5))
    p = M.gemm_plan(N, n, n, n, M.KernelConfig(thread_count=2, gemm_strategy="packed", gemm_panel_columns=5,
                                               gemm_micro_columns=4))
    assert d.strategy == "direct" and d.reason == "forced_direct"
    assert p.strategy == "packed" and p.reason == "forced_packed"
    assert p.panel_columns == 5 and p.micro_columns == 4
    assert p.packed_elements_per_worker == n * 5 and d.packed_elements_per_worker == 0
    assert d.workers == 2                                      # min(thread_count, cld(23, 5))
    assert d.network == p.network == "standard"               # packed and direct: acc + x*y


@pytest.mark.parametrize("N", [2, 3, 4])
def test_near_square_auto_policy(N):
    # runtests.jl:452-480 (profile: panel 16, micro 2, packed crossover 8 / 512, 1 thread)
    fused = N == 3
    cfg = M.KernelConfig(thread_count=1, gemm_panel_columns=16, gemm_micro_columns=2, gemm_packed_crossover=8)
    square, tall, reduction = (M.gemm_plan(N, *s, cfg) for s in ((64, 64, 64), (4096, 64, 64), (32, 16, 32)))
    assert (square.strategy, square.reason) == ("packed", "auto_above_crossover")
    assert tall.strategy == ("fused" if fused else "direct")
    assert tall.reason == ("auto_fused_direct" if fused else "auto_outside_calibrated_shape")
    assert reduction.strategy == ("fused" if fused else "direct")
    assert reduction.reason == ("auto_fused_direct" if fused else "auto_reduction_too_small")
    below = M.gemm_plan(N, 256, 256, 256, M.KernelConfig(thread_count=1, gemm_panel_columns=16, gemm_micro_columns=2,
                                                          gemm_packed_crossover=512))
    assert below.strategy == ("fused" if fused else "direct")
    assert below.reason == ("auto_fused_direct" if fused else "auto_below_crossover")


def test_x3_auto_is_fused_and_fused_needs_x3():
    # runtests.jl:2760-2775, 2516-2518
    a = M.gemm_plan(3, 32, 32, 32, M.KernelConfig(thread_count=1))
    assert (a.strategy, a.reason, a.network) == ("fused", "auto_fused_direct", "mulacc_x3")
    assert M.gemm_plan(3, 32, 32, 32, M.KernelConfig(thread_count=1, gemm_strategy="direct")).strategy == "direct"
    f = M.gemm_plan(3, 32, 32, 32, M.KernelConfig(gemm_strategy="fused"))
    assert (f.reason, f.micro_columns) == ("forced_fused", 2)
    with pytest.raises(M.ArgumentError):
        M.gemm_plan(2, 8, 8, 8, M.KernelConfig(gemm_strategy="fused"))
    with pytest.raises(M.ArgumentError):
        M.gemm_plan(5, 1, 1, 1, M.KernelConfig(thread_count=1))
    with pytest.raises(M.ArgumentError):
        M.gemm_plan(2, 8, 8, 8, M.KernelConfig(gemm_micro_columns=3))


def test_default_panel_columns_follow_thread_count():
    # gemm.jl:1-13: base 48 / 32 / 24 / 16, three quarters (at least 12) from 8 threads
    for N, base in ((1, 48), (2, 32), (3, 24), (4, 16)):
        assert M.gemm_plan(N, 64, 64, 64, M.KernelConfig(thread_count=1)).panel_columns == base
        assert M.gemm_plan(N, 64, 64, 64, M.KernelConfig(thread_count=8)).panel_columns == max(12, base * 3 // 4)
    assert M.gemm_plan(4, 64, 64, 64).micro_columns == 2 and M.gemm_plan(2, 64, 64, 64).micro_columns == 4


def test_device_launch_fields():
    # x4 n = 4096: the automatic 32 x 32 tile, one launch of 128 x 128 workgroups
    p = M.gemm_plan(4, 4096, 4096, 4096)
    assert (p.tile, p.tile_rows, p.tile_cols, p.launches, p.workgroups, p.remainder_rows) == (1, 32, 32, 1, 128 * 128, 0)
    # x3 mulacc route with m mod 4 = 3: main launch over 728 rows, remainder launch over 3 rows (16 x 16 tiles)
    q = M.gemm_plan(3, 731, 70, 720)
    assert (q.network, q.launches, q.remainder_rows, q.tile) == ("mulacc_x3", 2, 3, 1)
    assert q.workgroups == (728 // 32 + 1) * (720 // 32 + 1) + 1 * 45
    # x2 above 512 tiles of 64 x 64: the 64 x 64 tile
    r = M.gemm_plan(2, 2048, 2048, 2048)
    assert (r.tile, r.tile_rows, r.tile_cols) == (3, 64, 64)
    # a forced 64 x 64 tile for x4 runs the 32 x 64 tile
    t = M.gemm_plan(4, 1000, 100, 1000, M.KernelConfig(gemm_tile=3))
    assert (t.tile, t.tile_rows, t.tile_cols) == (2, 32, 64)
    z = M.gemm_plan(4, 0, 5, 7)
    assert z.launches == 0 and z.workgroups == 0


@pytest.mark.parametrize("N,block", [(1, 16), (2, 16), (3, 12), (4, 8)])
def test_ldlt_plan(N, block):
    # ldlt.jl:329-350, ldlt_tuning.jl (16 / 12 / 8), runtests.jl:1130-1139
    assert M.ldlt_plan(N, 511) == M.LDLTPlan("unblocked", "auto_below_crossover", 1)
    assert M.ldlt_plan(N, 512) == M.LDLTPlan("blocked", "auto_above_crossover", block)
    assert M.ldlt_plan(N, 100, M.KernelConfig(ldlt_strategy="unblocked")) == M.LDLTPlan("unblocked", "forced_unblocked", 1)
    p = M.ldlt_plan(N, 40, M.KernelConfig(thread_count=2, ldlt_strategy="blocked", ldlt_block=5, ldlt_blocked_crossover=1))
    assert p == M.LDLTPlan("blocked", "forced_blocked", 5)
    assert M.ldlt_plan(N, 40, M.KernelConfig(ldlt_strategy="blocked", ldlt_block=1)).block_size == 2
    assert M.ldlt_plan(N, 4096, M.KernelConfig(ldlt_block=1)).strategy == "unblocked"   # block 1: never blocked
    with pytest.raises(M.ArgumentError):
        M.ldlt_plan(N, 10, M.KernelConfig(ldlt_strategy="bogus"))


def test_sytrf_noalloc_rejects_what_does_not_fit_before_touching_anything():
    """mfla_sytrf_noalloc checks the static scratch size before it uses the
    handle or any pointer: limbs * n * (block + 2) doubles + 256 bytes must fit
    in MFLA_NOALLOC_SCRATCH_BYTES (8 MiB). Float64x4 (blocked, block 8): n =
    26213 fits (the null handle is reported next), n = 26214 does not; the
    unblocked route needs only the column maxima."""
    import ctypes

    from mfla import _lib

    lib = _lib.load()
    fake = 0x1000                       # never dereferenced: every call below returns before any access
    info = ctypes.c_int64(0)
    om = (ctypes.c_double * 4)()

    def call(N, n, strategy=0):
        cfg = _lib.MFConfig()
        lib.mfla_config_default(ctypes.byref(cfg))
        cfg.ldlt_strategy = strategy
        A = _lib.MFMatrix(fake, n, n, n, n * n)
        d = _lib.MFMatrix(fake, n, 1, n, n)
        return lib.mfla_sytrf_noalloc(None, N, A, d, ctypes.c_void_p(fake), ctypes.c_void_p(fake), om, ctypes.byref(info),
                                      ctypes.byref(cfg))

    assert call(4, 26214) == _lib.MFLA_ERR_WORKSPACE
    assert "device scratch" in _lib.last_error()
    assert call(4, 26213) == _lib.MFLA_ERR_ARGUMENT and "null handle" in _lib.last_error()
    # x1 (block 16): n = 58251 fits, 58252 does not
    assert call(1, 58252) == _lib.MFLA_ERR_WORKSPACE
    assert call(1, 58251) == _lib.MFLA_ERR_ARGUMENT
    # forced unblocked: the column maxima and their fold only
    assert call(4, 262136, strategy=1) == _lib.MFLA_ERR_ARGUMENT
    assert call(4, 262137, strategy=1) == _lib.MFLA_ERR_WORKSPACE
