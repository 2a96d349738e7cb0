"""tns_dgetrf_noalloc (the LU whose scratch is a file-scope device array, for
callers that cannot allocate) against tns_dgetrf on the same device inputs:
the same pivots, factors and first zero pivot, bit for bit, on two calls in
a row with the second smaller (the shared scratch keeps the first call's
contents and must be cleared), and its size and argument refusals."""
import ctypes as C

import numpy as np
import pytest

import csharp_linearalgebra_amd as T
from csharp_linearalgebra_amd import _lib as L

pytestmark = pytest.mark.gpu


def to_device(x):
    import torch

    t = torch.from_numpy(np.ascontiguousarray(np.asfortranarray(x).T)).cuda()
    torch.cuda.synchronize()  # the library runs on its own stream
    return T.DeviceMatrix.from_torch(t)


def noalloc(ws, d, nb=0):
    import torch

    k = min(d.rows, d.cols)
    piv = torch.full((max(k, 1),), -9, dtype=torch.int32, device="cuda")
    info = torch.full((1,), -9, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()  # the fills run on torch's stream, the LU on the library's
    rc = L.load().tns_dgetrf_noalloc(ws.handle, d.rows, d.cols, d.ptr, d.ld, nb, piv.data_ptr(), info.data_ptr())
    ws.synchronize()
    return rc, piv[:k].cpu().numpy(), int(info.item())


def compare(ws, a, nb=0):
    ref_d, got_d = to_device(a), to_device(a)
    piv_ref, sing_ref = ws.factor_lu_device(ref_d, block_size=nb)
    rc, piv, info = noalloc(ws, got_d, nb)
    assert rc == L.TNS_OK
    assert info == sing_ref
    assert np.array_equal(piv, piv_ref)
    assert np.array_equal(got_d.owner.cpu().numpy(), ref_d.owner.cpu().numpy())


def test_two_calls_in_a_row_match_the_allocating_lu(ws):
    r = np.random.default_rng(11)
    compare(ws, r.uniform(-1, 1, (3000, 3000)))
    compare(ws, r.uniform(-1, 1, (700, 700)))  # smaller: stale scratch from the first call
    compare(ws, r.uniform(-1, 1, (900, 260)), nb=32)
    compare(ws, r.uniform(-1, 1, (130, 400)))


def test_a_zero_column_is_reported(ws):
    a = np.random.default_rng(12).uniform(-1, 1, (300, 300))
    a[:, 77] = 0.0
    compare(ws, a)
    _, _, info = noalloc(ws, to_device(a))
    assert info == 77


def test_refusals(ws):
    lib = L.load()
    import torch

    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    piv = torch.zeros(8, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    # Past the scratch it holds: refused before anything is read.
    assert lib.tns_dgetrf_noalloc(ws.handle, 40000, 40000, None, 40000, 0, None, info.data_ptr()) == L.TNS_ERANGE
    assert b"scratch" in lib.tns_last_error(ws.handle)
    d = to_device(np.eye(8))
    assert lib.tns_dgetrf_noalloc(ws.handle, 8, 8, d.ptr, 8, 128, piv.data_ptr(), info.data_ptr()) == L.TNS_EINVAL
    host_info = np.zeros(1, dtype=np.int32)
    assert lib.tns_dgetrf_noalloc(ws.handle, 8, 8, d.ptr, 8, 0, piv.data_ptr(),
                                  host_info.ctypes.data) == L.TNS_EINVAL
    a = np.eye(8)
    assert lib.tns_dgetrf_noalloc(ws.handle, 8, 8, a.ctypes.data, 8, 0, piv.data_ptr(),
                                  info.data_ptr()) == L.TNS_EINVAL
    # Empty: nothing to do, info -1.
    assert lib.tns_dgetrf_noalloc(ws.handle, 0, 5, None, 1, 0, None, info.data_ptr()) == L.TNS_OK
    ws.synchronize()
    assert int(info.item()) == -1
