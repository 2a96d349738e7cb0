"""The C ABI library loads and exports every entry point the public header
declares, with the ctypes binding covering each one. No device call."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tensile_hip.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"TNS_API\s+[\w\s\*]+?\b(tns_\w+)\s*\(", text)))


def test_header_declares_the_seam():
    names = declared()
    for must in ("tns_dgemm", "tns_zgemm", "tns_dgetrf", "tns_zgetrf", "tns_dgetrs", "tns_zgetrs", "tns_dtrsm",
                 "tns_dpanel", "tns_dnorm", "tns_dgecon", "tns_zgecon", "tns_dexpm", "tns_zexpm", "tns_dexpmv",
                 "tns_zexpmv", "tns_dexpmv_op", "tns_zexpmv_op"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import csharp_linearalgebra_amd as T

    lib = ctypes.CDLL(T.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    import csharp_linearalgebra_amd as T

    assert sorted(T.exported_symbols()) == declared()


def test_version_string():
    import csharp_linearalgebra_amd as T

    assert b"gfx950" in T.load().tns_version()


def test_context_creation_without_a_device_fails_cleanly():
    import csharp_linearalgebra_amd as T

    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("a device is present")
    except ImportError:
        pass
    with pytest.raises(T.DeviceError):
        T.Workspace(0)
