"""Mixed-precision solve: FP32 (and BF16-product) LU with FP64 iterative
refinement, falling back to the FP64 LU. The reference has no such path;
the result is pinned by its backward error (at most 1e-15, in the
ComplexLuTests.cs:418-439 form) and by the FP64 solve of the oracle
(Lu.cs:189-378) it must agree with to the conditioning."""
import numpy as np
import pytest

import oracle as O
from csharp_linearalgebra_amd import workloads as W

pytestmark = pytest.mark.gpu


def berr(a, x, b):
    r = b - a @ x
    den = np.abs(a).sum(1).max() * np.abs(x).max(0) + np.abs(b).max(0)
    return float((np.abs(r).max(0) / den).max())


def forward_ok(a, x, b, kappa):
    lu = O.getrf(a)
    ref = O.getrs(lu, b)
    rel = np.abs(x - ref).max() / np.abs(ref).max()
    return rel <= 50 * kappa * np.finfo(float).eps, rel


@pytest.mark.parametrize("n,kappa", [(256, 1e2), (500, 1e2), (384, 1e6)])
def test_fp32_refinement_reaches_fp64_backward_error(ws, n, kappa):
    a, b = W.conditioned(n, kappa, 8, seed=3)
    x, info = ws.solve_mixed(a, b, precision=32)
    assert info.first_error <= 1e-5, info
    assert not info.fell_back, info
    assert 1 <= info.iterations <= 30
    assert info.backward_error <= 1e-15
    assert berr(a, x, b) <= 2e-15
    ok, rel = forward_ok(a, x, b, kappa)
    assert ok, rel


def test_fp32_falls_back_when_refinement_stalls(ws):
    # kappa * u_fp32 >> 1: the residuals shrink by about 1% a step, so the
    # first FGMRES cycle gives up at its fourth step (the extrapolated steps
    # would cost far more than the FP64 solve).
    a, b = W.conditioned(256, 1e10, 4, seed=5)
    x, info = ws.solve_mixed(a, b, precision=32)
    assert info.fell_back
    assert info.iterations == 5, info
    assert info.backward_error <= 1e-15
    assert berr(a, x, b) <= 2e-15


def test_falls_back_when_the_solve_budget_runs_out(ws):
    a, b = W.conditioned(384, 1e6, 4, seed=3)
    x, info = ws.solve_mixed(a, b, precision=32, max_iterations=2)
    assert info.fell_back and info.iterations == 2, info
    assert berr(a, x, b) <= 2e-15
    # One solve only: no refinement at all.
    x, info = ws.solve_mixed(a, b, precision=32, max_iterations=1)
    assert info.fell_back and info.iterations == 1, info
    assert berr(a, x, b) <= 2e-15


def test_bf16_products_with_equilibration(ws):
    a, b = W.conditioned(320, 10.0, 4, seed=7)
    # Badly scaled rows and columns: equilibration must undo them.
    rs = 2.0 ** np.arange(-6, 6, 12 / 320)[:320]
    a = np.asfortranarray(a * rs[:, None] * rs[::-1][None, :])
    b = np.asfortranarray(b * rs[:, None])
    x, info = ws.solve_mixed(a, b, precision=16, max_iterations=60)
    assert info.backward_error <= 1e-15
    assert berr(a, x, b) <= 2e-15
    assert not info.fell_back, info


def test_tile_aligned_and_ragged_orders(ws):
    for n in (64, 127, 1024):
        a, b = W.conditioned(n, 1e3, 3, seed=n)
        x, info = ws.solve_mixed(a, b, precision=32)
        assert info.backward_error <= 1e-15, (n, info)
        assert berr(a, x, b) <= 2e-15


def test_device_workload_matches_host_generator():
    import torch

    a, b = W.conditioned(96, 1e6, 5, seed=11)
    ad, bd = W.conditioned_device(96, 1e6, 5, seed=11)
    torch.cuda.synchronize()
    assert np.abs(ad.cpu().numpy() - a).max() <= 1e-12
    assert np.array_equal(bd.cpu().numpy(), b)


def test_argument_checks(ws):
    import csharp_linearalgebra_amd as T

    a = np.eye(4)
    b = np.ones((4, 1))
    with pytest.raises(T.ArgumentError):
        ws.solve_mixed(a, b, precision=8)
    with pytest.raises(T.ArgumentOutOfRangeError):
        ws.solve_mixed(a, b, tol=0.0)
    with pytest.raises(T.InvalidOperationError):
        ws.solve_mixed(np.zeros((4, 4)), b)


@pytest.mark.parametrize("precision", [32, 16])
@pytest.mark.parametrize("m,n,k", [(64, 64, 64), (100, 37, 129), (256, 192, 64), (17, 5, 3)])
def test_single_precision_products(ws, rng, precision, m, n, k):
    import csharp_linearalgebra_amd as T
    from tests._raw import addr

    lib = T.load()
    a = np.asfortranarray(rng.uniform(-1, 1, (m, k)).astype(np.float32))
    b = np.asfortranarray(rng.uniform(-1, 1, (k, n)).astype(np.float32))
    c0 = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(np.float32))
    c = c0.copy(order="F")
    assert lib.tns_sgemm(ws.handle, precision, m, n, k, -1.0, addr(a), m, addr(b), k, 1.0, addr(c), m) == 0
    if precision == 16:
        # Operands rounded to bfloat16 (nearest, ties to even), products
        # and sums in FP32.
        def bf(x):
            u = x.astype(np.float32).view(np.uint32).astype(np.uint64)
            u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
            return u.astype(np.uint32).view(np.float32).astype(np.float64)
        ref = c0 - bf(a) @ bf(b)
    else:
        ref = c0.astype(np.float64) - a.astype(np.float64) @ b.astype(np.float64)
    bound = (k + 2) * np.finfo(np.float32).eps * (np.abs(a).astype(np.float64) @ np.abs(b) + np.abs(c0))
    assert np.all(np.abs(c - ref) <= 2 * bound + 1e-30)


@pytest.mark.parametrize("m,n,k", [(64, 64, 64), (100, 37, 129), (256, 192, 300)])
def test_three_part_bf16_products(ws, rng, m, n, k):
    """Precision 48: operands split x = hi + lo into two bfloat16 parts and
    the product summed as hi*hi + hi*lo + lo*hi in FP32. The error per
    product is about 2^-16 of |a||b| (the dropped lo*lo and lo's rounding),
    far inside one-part BF16's 2^-8: checked against the FP64 product, and
    against the one-part error on the same data (a kernel that dropped a
    cross term or rounded only once would fail the second bound)."""
    import csharp_linearalgebra_amd as T
    from tests._raw import addr

    lib = T.load()
    a = np.asfortranarray(rng.uniform(-1, 1, (m, k)).astype(np.float32))
    b = np.asfortranarray(rng.uniform(-1, 1, (k, n)).astype(np.float32))
    c0 = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(np.float32))
    out = {}
    for prec in (48, 16):
        c = c0.copy(order="F")
        assert lib.tns_sgemm(ws.handle, prec, m, n, k, -1.0, addr(a), m, addr(b), k, 1.0, addr(c), m) == 0
        out[prec] = c.astype(np.float64)
    ref = c0.astype(np.float64) - a.astype(np.float64) @ b.astype(np.float64)
    mag = np.abs(a).astype(np.float64) @ np.abs(b) + np.abs(c0)
    err3 = np.abs(out[48] - ref)
    assert np.all(err3 <= (k * 2.0 ** -15 + (k + 2) * np.finfo(np.float32).eps) * mag)
    err1 = np.abs(out[16] - ref)
    assert err3.max() * 20 < err1.max()


@pytest.mark.parametrize("n", [48, 200, 700, 1100, 2600])
def test_single_precision_lu_matches_lapack_sgetrf(ws, rng, n):
    import scipy.linalg as sl

    import csharp_linearalgebra_amd as T
    from tests._raw import addr

    lib = T.load()
    a = np.asfortranarray(rng.uniform(-1, 1, (n, n)).astype(np.float32))
    f = a.copy(order="F")
    piv = np.zeros(n, dtype=np.int32)
    sc = C_int()
    import ctypes
    assert lib.tns_sgetrf(ws.handle, 32, n, addr(f), n, addr(piv), ctypes.byref(sc)) == 0
    lu, ref_piv = sl.lu_factor(a)
    # The pivot rule is LAPACK's; in FP32 the roundings of the blocked and
    # the recursive (sgetrf2) orders differ, so after the first near-tie the
    # sequences may part. What partial pivoting guarantees for any order:
    # every multiplier is at most 1 in magnitude, and the pivots agree until
    # the first near-tie.
    first = int(np.argmax(piv != ref_piv)) if np.any(piv != ref_piv) else n
    assert first >= min(n, 16)
    assert np.abs(np.tril(f, -1)).max() <= 1.0
    # Backward error of the FP32 factors.
    l = np.tril(f.astype(np.float64), -1) + np.eye(n)
    u = np.triu(f.astype(np.float64))
    pa = a.astype(np.float64).copy()
    for k, p in enumerate(piv):
        pa[[k, p]] = pa[[p, k]]
    assert np.abs(pa - l @ u).sum(0).max() / np.abs(a).sum(0).max() <= n * np.finfo(np.float32).eps
    b = np.asfortranarray(rng.uniform(-1, 1, (n, 3)).astype(np.float32))
    x = b.copy(order="F")
    assert lib.tns_sgetrs(ws.handle, 32, n, 3, addr(f), n, addr(piv), addr(x), n) == 0
    r = b.astype(np.float64) - a.astype(np.float64) @ x.astype(np.float64)
    assert np.abs(r).max() / (np.abs(a).sum(1).max() * np.abs(x).max()) <= n * np.finfo(np.float32).eps


def C_int():
    import ctypes

    return ctypes.c_int()


@pytest.mark.parametrize("precision", [32, 16])
def test_order_8192_outer_steps(ws, precision):
    # n = 8192 takes the block-local LU's 256-column outer steps; the
    # matrix is generated and the residual checked on the device.
    import torch

    from csharp_linearalgebra_amd import DeviceMatrix

    n = 8192
    a, b = W.conditioned_device(n, 1e2, 4, seed=2)
    at, bt = a.T.contiguous(), b.T.contiguous()  # column-major storage
    dA, dB = DeviceMatrix.from_torch(at), DeviceMatrix.from_torch(bt)
    dX = DeviceMatrix.from_torch(torch.empty_like(bt))
    torch.cuda.synchronize()  # the library's stream does not wait on torch's
    _, info = ws.solve_mixed(dA, dB, precision=precision, out=dX)
    ws.synchronize()
    assert not info.fell_back, info
    assert info.backward_error <= 1e-15, info
    x = dX.owner.T  # the (cols, rows) tensor behind the column-major matrix
    r = b - a @ x
    den = a.abs().sum(1).max() * x.abs().max(0).values + b.abs().max(0).values
    assert float((r.abs().max(0).values / den).max()) <= 2e-15
