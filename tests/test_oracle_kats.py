"""The reference's known-answer tests (test/adversarial.jl, test/runtests.jl)
run against the oracle on the CPU: pins the restated factorizations, solves,
residuals and backward error to MFLA's closed-form answers."""
import pytest

from backends import OracleBackend
from kat_cases import run_all


@pytest.mark.parametrize("N", [2, 3, 4])
def test_reference_kats_on_oracle(oracle, N):
    res = run_all(OracleBackend(oracle), N)
    failed = [m for ok, m in res if not ok]
    assert not failed, failed
    assert len(res) >= 45
