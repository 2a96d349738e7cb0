"""Loads libtensile_hip.so and declares its C ABI (include/tensile_hip.h).

There is no fallback: if the library is missing or does not load, importing
the backend raises, so nothing can run on a silent substitute.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("TENSILE_HIP_LIBRARY", os.path.join(HERE, "libtensile_hip.so"))

TNS_OK = 0
TNS_EINVAL = 1
TNS_ERANGE = 2
TNS_ESTATE = 3
TNS_ENOMEM = 4
TNS_EDISPOSED = 5
TNS_EDEVICE = 6
TNS_ECALLBACK = 7


class Complex(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class LuInfo(C.Structure):
    _fields_ = [("singular_column", C.c_int), ("smallest_pivot", C.c_double), ("largest_pivot", C.c_double)]


class NormEst(C.Structure):
    _fields_ = [("value", C.c_double), ("iterations", C.c_int), ("products", C.c_int)]


class ExpmDiag(C.Structure):
    _fields_ = [("degree", C.c_int), ("squarings", C.c_int)]


class MixedInfo(C.Structure):
    _fields_ = [("precision", C.c_int), ("iterations", C.c_int), ("fell_back", C.c_int),
                ("backward_error", C.c_double), ("factor_ms", C.c_double), ("refine_ms", C.c_double),
                ("fallback_ms", C.c_double), ("first_error", C.c_double)]


class LuPhases(C.Structure):
    _fields_ = [("panel_ms", C.c_double), ("swaps_ms", C.c_double), ("trsm_ms", C.c_double),
                ("gemm_ms", C.c_double), ("total_ms", C.c_double), ("block_steps", C.c_int)]


class ExpmPhases(C.Structure):
    _fields_ = [("powers_ms", C.c_double), ("select_ms", C.c_double), ("pade_ms", C.c_double),
                ("lu_ms", C.c_double), ("solve_ms", C.c_double), ("squaring_ms", C.c_double),
                ("total_ms", C.c_double)]


class ExpmvDiag(C.Structure):
    _fields_ = [("degree", C.c_int), ("scaling", C.c_int), ("applications", C.c_int), ("early_exits", C.c_int)]


DAPPLY = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int)
ZAPPLY = DAPPLY

# name -> (restype, argtypes). Pointers to matrices are c_void_p: the caller
# passes host addresses (numpy) or device addresses (torch) alike.
_P, _I, _D, _CH, _SZ = C.c_void_p, C.c_int, C.c_double, C.c_char, C.c_size_t
SIGNATURES = {
    "tns_context_create": (_I, [_I, C.POINTER(C.c_void_p)]),
    "tns_context_destroy": (_I, [_P]),
    "tns_last_error": (C.c_char_p, [_P]),
    "tns_context_set_limit": (_I, [_P, _SZ]),
    "tns_context_set_retention": (_I, [_P, _SZ]),
    "tns_context_memory": (_I, [_P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "tns_context_stream": (_I, [_P, C.POINTER(C.c_void_p)]),
    "tns_synchronize": (_I, [_P]),
    "tns_version": (C.c_char_p, []),
    "tns_gemm_set_config": (_I, [_I, _I]),
    "tns_gemm_config_count": (_I, [_I]),
    "tns_dgemm": (_I, [_P, _I, _I, _I, _D, _P, _I, _P, _I, _D, _P, _I]),
    "tns_zgemm": (_I, [_P, _I, _I, _I, Complex, _P, _I, _P, _I, Complex, _P, _I]),
    "tns_dgetrf": (_I, [_P, _I, _I, _P, _I, _I, _P, C.POINTER(LuInfo)]),
    "tns_zgetrf": (_I, [_P, _I, _I, _P, _I, _I, _P, C.POINTER(LuInfo)]),
    "tns_dgetrf_timed": (_I, [_P, _I, _I, _P, _I, _I, _P, C.POINTER(LuInfo), C.POINTER(LuPhases)]),
    "tns_dgetrf_noalloc": (_I, [_P, _I, _I, _P, _I, _I, _P, _P]),
    "tns_zgetrf_timed": (_I, [_P, _I, _I, _P, _I, _I, _P, C.POINTER(LuInfo), C.POINTER(LuPhases)]),
    "tns_dgetrs": (_I, [_P, _I, _I, _I, _P, _I, _P, _I, _P, _I]),
    "tns_zgetrs": (_I, [_P, _I, _I, _I, _P, _I, _P, _I, _P, _I]),
    "tns_dtrsm": (_I, [_P, _I, _I, _I, _P, _I, _P, _I]),
    "tns_dpanel": (_I, [_P, _I, _I, _I, _P, _I, _P, _I, _P, _I]),
    "tns_zpanel": (_I, [_P, _I, _I, _I, _P, _I, _P, _I, _P, _I]),
    "tns_dnorm": (_I, [_P, _CH, _I, _I, _P, _I, C.POINTER(_D)]),
    "tns_znorm": (_I, [_P, _CH, _I, _I, _P, _I, C.POINTER(_D)]),
    "tns_dnormest_power": (_I, [_P, _I, _P, _I, _I, _I, _I, _I, C.POINTER(NormEst)]),
    "tns_znormest_power": (_I, [_P, _I, _P, _I, _I, _I, _I, _I, C.POINTER(NormEst)]),
    "tns_dgecon": (_I, [_P, _I, _P, _I, _P, _I, _D, _I, C.POINTER(_D)]),
    "tns_zgecon": (_I, [_P, _I, _P, _I, _P, _I, _D, _I, C.POINTER(_D)]),
    "tns_dgesv_mixed": (_I, [_P, _I, _I, _P, _I, _P, _I, _P, _I, _I, _D, _I, C.POINTER(MixedInfo)]),
    "tns_sgemm": (_I, [_P, _I, _I, _I, _I, C.c_float, _P, _I, _P, _I, C.c_float, _P, _I]),
    "tns_sgetrf": (_I, [_P, _I, _I, _P, _I, _P, C.POINTER(_I)]),
    "tns_sgetrs": (_I, [_P, _I, _I, _I, _P, _I, _P, _P, _I]),
    "tns_dexpmv_grad": (_I, [_P, _I, _I, _P, _I, _P, _I, _D, _P, _I, _P, _I, _P, _I, C.POINTER(ExpmvDiag)]),
    "tns_zexpmv_grad": (_I, [_P, _I, _I, _P, _I, _P, _I, _D, _P, _I, _P, _I, _P, _I, C.POINTER(ExpmvDiag)]),
    "tns_zbatch_create": (_I, [_P, _I, _P, _P, _P, _P, _P, _I, _I, C.POINTER(C.c_void_p)]),
    "tns_zbatch_solve": (_I, [_P, _P]),
    "tns_zbatch_rcond": (_I, [_P, _I, _P]),
    "tns_zbatch_pivots": (_I, [_P, _I, _P]),
    "tns_zbatch_info": (_I, [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_D)]),
    "tns_zbatch_destroy": (_I, [_P]),
    "tns_dexpm": (_I, [_P, _I, _P, _I, _P, _I, C.POINTER(ExpmDiag)]),
    "tns_zexpm": (_I, [_P, _I, _P, _I, _P, _I, C.POINTER(ExpmDiag)]),
    "tns_dexpm_timed": (_I, [_P, _I, _P, _I, _P, _I, C.POINTER(ExpmDiag), C.POINTER(ExpmPhases)]),
    "tns_zexpm_timed": (_I, [_P, _I, _P, _I, _P, _I, C.POINTER(ExpmDiag), C.POINTER(ExpmPhases)]),
    "tns_dexpmv": (_I, [_P, _I, _I, _P, _I, _P, _I, _D, _P, _I, C.POINTER(ExpmvDiag)]),
    "tns_zexpmv": (_I, [_P, _I, _I, _P, _I, _P, _I, _D, _P, _I, C.POINTER(ExpmvDiag)]),
    "tns_dexpmv_op": (_I, [_P, _I, DAPPLY, _P, _I, _I, _P, _I, _D, _D, _P, _I, C.POINTER(ExpmvDiag)]),
    "tns_zexpmv_op": (_I, [_P, _I, ZAPPLY, _P, _I, _I, _P, _I, _D, _D, _P, _I, C.POINTER(ExpmvDiag)]),
}

_lib = None


def _share_torch_runtime():
    """PyTorch's ROCm wheel carries its own libamdhip64 and asks for it by a
    name ("libamdhip64.so") that differs from the soname this library asks
    for ("libamdhip64.so.7"). Loaded in the other order the process would
    hold two HIP runtimes, and the second one finds no device; with torch's
    loaded first the dynamic linker resolves this library's request to it by
    soname, so both share one runtime, one device context and one address
    space of device pointers."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load():
    """The loaded library; raises OSError if it is not built."""
    global _lib
    if _lib is None:
        _share_torch_runtime()
        if not os.path.exists(LIB_PATH):
            raise OSError(
                f"libtensile_hip.so not found at {LIB_PATH}; build it with "
                "`make -C csharp_linearalgebra_amd` (or __graft_entry__.build())")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def exported_symbols():
    return sorted(SIGNATURES)
