"""calibrate_gemm on the device: every tile candidate is bitwise equal to the
automatic choice (the acceptance gate of src/calibration.jl:178-335), the
chosen tile is one of the candidates or the automatic one, and the profile
applies on the device it was made on."""
import pytest

import mfla as M

from mfutil import HELD_OFF_GPU

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N", [2, pytest.param(3, marks=HELD_OFF_GPU), 4])
def test_calibrate_gemm(N):
    prof, meas = M.calibrate_gemm(N, sizes=(192, 320), samples=1)
    assert prof.source == "calibrated" and prof.limbs == N
    assert prof.tile in (-1, 0, 1, 2, 3)
    assert all(m["bitwise_equal"] for m in meas if m["tile"] >= 0)
    assert {m["size"] for m in meas} == {192, 320}
    assert M.profile_compatible(prof)
    cfg = M.with_gemm_profile(None, prof)
    assert cfg.gemm_tile == prof.tile
    assert M.GemmProfile.from_json(prof.to_json()) == prof
