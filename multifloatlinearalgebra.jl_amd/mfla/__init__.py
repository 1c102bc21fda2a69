"""mfla -- MI355X (gfx950) device backend for MultiFloatLinearAlgebra.jl's API.

Python host mirror of the reference's public surface; every numerical
operation runs in libmfla_hip.so on the GPU (see include/mfla_hip.h).
"""
from ._lib import LIB_PATH, load  # noqa: F401
from .api import (  # noqa: F401
    ArgumentError, ConvergenceError, DimensionMismatch, add_mixed_, DomainError, KernelConfig, MFCholesky, MFLDLT, MFLU, MFWorkspace, PosDefException,
    SingularException, arith, cholesky_, ensure_workspace_capacity_, from_interleaved, gemm_, gemmt_, gemv_, host_arith, ldiv_, ldlt_,
    lu_, normwise_backward_error, refinement_correction_, residual_, residual_mixed_, solve, symv_, synchronize, syrk_, to_interleaved,
    trsm_, trsv_, workspace_capacity, trmm_, mfdot, syrk_packed_, capabilities, BoundsError, dgemm_, dpotrf_inverse_, dgetrf_,
    MFQR, rrqr_, qr_, apply_q_, solve_r_, GemmPlan, LDLTPlan, gemm_plan, ldlt_plan, ldlt_noalloc_,
)
from .factor import (  # noqa: F401
    factor_blocks, factor_diagnostics, factor_inertia, factor_kind, factor_matrix, factor_permutation, factor_pivots,
    factor_precision, factor_provider, factor_state, factor_status, issuccess, numerical_rank, factor_rdiag,
)
from .calibration import (  # noqa: F401
    GemmProfile, calibrate_gemm, default_gemm_profile, machine_fingerprint, profile_compatible, with_gemm_profile,
)
