"""MI355X-native backend for the dense hot path of Tensile
(xfmrexpert/CSharp_LinearAlgebra): FP64-MFMA GEMM and 4M complex GEMM, LU
with partial pivoting and its solves, normest1/gecon, Pade scaling-and-
squaring Expm and Taylor Expmv, behind the C ABI in include/tensile_hip.h.

The compute lives in libtensile_hip.so (HIP, gfx950). This package is the
Python mirror of the reference's public surface over that ABI; it has no
fallback path: without the library every call raises.
"""
from ._lib import LIB_PATH, exported_symbols, load  # noqa: F401
from .tensile import (  # noqa: F401
    AllocationLimitError,
    ArgumentError,
    ArgumentOutOfRangeError,
    BatchedSystems,
    DenseDeviceOperator,
    DeviceError,
    DeviceLinearOperator,
    DeviceMatrix,
    ExpmDiagnostics,
    ExpmvDiagnostics,
    ExpmvFunction,
    ILinearOperator,
    InvalidOperationError,
    LuDecomposition,
    LuPhaseTimings,
    ExpmPhaseTimings,
    MixedInfo,
    NormEstimateResult,
    ObjectDisposedError,
    OperatorError,
    TensileError,
    Workspace,
    estimate_one_norm,
    expm,
    expmv,
    factor_lu,
    frobenius_norm,
    infinity_norm,
    multiply,
    multiply_into,
    one_norm,
    solve,
)

TRI_LOWER_UNIT, TRI_UPPER, TRI_LOWER, TRI_UPPER_T, TRI_LOWER_T, TRI_LOWER_UNIT_T = range(6)
