"""FP64-MFMA GEMM and 4M complex GEMM against the oracle, on the shapes and
scalar cases of the reference's GEMM contracts (GemmTests.cs:31-129,
ComplexMultiplyTests.cs:40-315). Bound: elementwise
|C - C_ref| <= (k+2) eps (|alpha| |A||B| + |beta||C|): the k eps (|A||B|)
GEMM bound with room for the oracle's own rounding."""
import numpy as np
import pytest

import oracle as O
from tests._raw import addr, cplx, padded

pytestmark = pytest.mark.gpu
EPS = np.finfo(float).eps

SHAPES = [(1, 1, 1), (1, 1, 64), (8, 6, 4), (16, 8, 16), (17, 9, 5), (31, 33, 7), (48, 48, 48), (64, 64, 64),
          (65, 63, 66), (100, 37, 129), (129, 130, 2), (300, 257, 513), (257, 1031, 96)]


def bound(a, b, c, alpha, beta, k):
    return (k + 2) * EPS * (abs(alpha) * (np.abs(a) @ np.abs(b)) + abs(beta) * np.abs(c)) + 1e-300


def test_mfma_block_layout(ws):
    # A 16x16x4 product with distinct entries pins every lane's row/column.
    a = np.arange(16 * 4, dtype=float).reshape(16, 4, order="F") + 1
    b = (np.arange(4 * 16, dtype=float).reshape(4, 16, order="F") + 1) * 0.5
    c = ws.multiply(a, b)
    assert np.array_equal(c, a @ b)


@pytest.mark.parametrize("m,n,k", SHAPES)
def test_dgemm_matches_oracle(ws, rng, m, n, k):
    a = rng.uniform(-0.5, 0.5, (m, k))
    b = rng.uniform(-0.5, 0.5, (k, n))
    c = ws.multiply(a, b)
    ref = O.gemm(a, b)
    assert np.all(np.abs(c - ref) <= 2 * bound(a, b, 0 * ref, 1, 0, k))


@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (1.0, 1.0), (-1.0, 1.0), (0.0, 1.0), (0.0, 0.0), (2.5, -0.75)])
def test_dgemm_scalars(ws, rng, alpha, beta):
    a = rng.uniform(-0.5, 0.5, (37, 41))
    b = rng.uniform(-0.5, 0.5, (41, 29))
    c0 = np.asfortranarray(rng.uniform(-0.5, 0.5, (37, 29)))
    c = c0.copy(order="F")
    ws.multiply(a, b, c, alpha, beta)
    ref = O.gemm(a, b, c0, alpha, beta)
    assert np.all(np.abs(c - ref) <= 2 * bound(a, b, c0, alpha, beta, 41))


def test_beta_zero_overwrites_nan(ws, rng):
    a = rng.uniform(-0.5, 0.5, (33, 17))
    b = rng.uniform(-0.5, 0.5, (17, 19))
    c = np.full((33, 19), np.nan, order="F")
    ws.multiply(a, b, c, 1.0, 0.0)
    assert np.all(np.isfinite(c))
    c = np.full((33, 19), np.nan, order="F")
    ws.multiply(a[:, :0], b[:0, :], c, 1.0, 0.0)  # k == 0 still scales (here: zeroes) C
    assert np.all(c == 0.0)


@pytest.mark.parametrize("pad", [1, 3, 17])
def test_padding_rows_are_never_written(ws, rng, pad):
    import csharp_linearalgebra_amd as T

    lib = T.load()
    m, n, k = 35, 27, 19
    abuf, a = padded(m, k, m + pad)
    bbuf, b = padded(k, n, k + pad)
    cbuf, c = padded(m, n, m + pad)
    a[...] = rng.uniform(-0.5, 0.5, (m, k))
    b[...] = rng.uniform(-0.5, 0.5, (k, n))
    c[...] = rng.uniform(-0.5, 0.5, (m, n))
    c0 = c.copy()
    rc = lib.tns_dgemm(ws.handle, m, n, k, 1.0, addr(abuf), m + pad, addr(bbuf), k + pad, 0.5, addr(cbuf), m + pad)
    assert rc == 0
    assert np.all(np.isnan(cbuf[m:, :]))
    ref = O.gemm(a, b, c0, 1.0, 0.5)
    assert np.all(np.abs(c - ref) <= 2 * bound(a, b, c0, 1, 0.5, k))


def test_empty_destination_returns(ws):
    import csharp_linearalgebra_amd as T

    lib = T.load()
    # 0 x 2^31-1 nominal columns: nothing is touched (KernelEntry.cs:41-44).
    assert lib.tns_dgemm(ws.handle, 0, 2147483647, 5, 1.0, None, 1, None, 5, 0.0, None, 1) == 0


def test_shape_disagreement_is_an_argument_error(ws):
    import csharp_linearalgebra_amd as T

    with pytest.raises(T.ArgumentError):
        ws.multiply(np.ones((3, 4)), np.ones((5, 2)))


@pytest.mark.parametrize("m,n,k", SHAPES)
def test_zgemm_4m_matches_oracle(ws, rng, m, n, k):
    a = rng.uniform(-0.5, 0.5, (m, k)) + 1j * rng.uniform(-0.5, 0.5, (m, k))
    b = rng.uniform(-0.5, 0.5, (k, n)) + 1j * rng.uniform(-0.5, 0.5, (k, n))
    c = ws.multiply(a, b)
    ref = O.gemm(a, b)
    assert np.all(np.abs(c - ref) <= 4 * bound(a, b, 0 * ref, 1, 0, k))


@pytest.mark.parametrize("alpha,beta", [(1, 0), (1, 1), (2.5, 0), (0.5 - 1.5j, 0), (1, -0.25 + 2j), (0.3 + 0.7j, 0.9 - 0.1j),
                                        (0, 1), (0, 0)])
def test_zgemm_scalars(ws, rng, alpha, beta):
    a = rng.uniform(-0.5, 0.5, (23, 31)) + 1j * rng.uniform(-0.5, 0.5, (23, 31))
    b = rng.uniform(-0.5, 0.5, (31, 19)) + 1j * rng.uniform(-0.5, 0.5, (31, 19))
    c0 = np.asfortranarray(rng.uniform(-0.5, 0.5, (23, 19)) + 1j * rng.uniform(-0.5, 0.5, (23, 19)))
    c = c0.copy(order="F")
    ws.multiply(a, b, c, complex(alpha), complex(beta))
    ref = O.gemm(a, b, c0, complex(alpha), complex(beta))
    assert np.all(np.abs(c - ref) <= 4 * bound(a, b, c0, alpha, beta, 31))


def test_small_imaginary_part_keeps_relative_accuracy(ws, rng):
    # ComplexMultiplyTests.cs:184-210: Im(AB) tiny against Re(AB); 4M keeps
    # it to working accuracy, the 3M shortcut (oracle) loses it.
    n = 64
    for rho in (1e3, 1e6, 1e9, 1e12):
        a = rng.uniform(0.5, 1.0, (n, n)) + 1j * rng.uniform(-1, 1, (n, n)) / rho
        b = rng.uniform(0.5, 1.0, (n, n)) + 1j * rng.uniform(-1, 1, (n, n)) / rho
        c = ws.multiply(a, b)
        exact_im = a.real @ b.imag + a.imag @ b.real
        scale = np.abs(a.real) @ np.abs(b.imag) + np.abs(a.imag) @ np.abs(b.real)
        assert np.max(np.abs(c.imag - exact_im) / scale) <= (n + 2) * EPS * 4
        c3 = O.zgemm_3m(a, b)
        if rho >= 1e6:
            assert np.max(np.abs(c3.imag - exact_im) / scale) > (n + 2) * EPS * 4


def test_real_operands_give_the_real_product(ws, rng):
    a = rng.uniform(-1, 1, (40, 30))
    b = rng.uniform(-1, 1, (30, 20))
    z = ws.multiply(a.astype(complex), b.astype(complex))
    r = ws.multiply(a, b)
    assert np.all(z.imag == 0.0)
    assert np.max(np.abs(z.real - r)) <= 32 * EPS


def test_large_square_gemm(ws, rng):
    n = 2048
    a = rng.uniform(-1, 1, (n, n))
    b = rng.uniform(-1, 1, (n, n))
    c = ws.multiply(a, b)
    ref = O.gemm(a, b)
    assert np.all(np.abs(c - ref) <= 2 * bound(a, b, 0 * ref, 1, 0, n))


CONFIG_SHAPES = [(17, 9, 5), (65, 63, 66), (256, 128, 64), (300, 257, 513), (512, 256, 48)]


@pytest.mark.parametrize("complex_", [False, True])
def test_every_tile_config_matches_oracle(ws, rng, complex_):
    # Each tile shape the heuristic can pick, forced in turn, on ragged and
    # on tile-aligned shapes (the latter take the paired 16-byte loads).
    import csharp_linearalgebra_amd as T

    lib = T.load()
    try:
        for cfg in range(lib.tns_gemm_config_count(int(complex_))):
            lib.tns_gemm_set_config(-1 if complex_ else cfg, cfg if complex_ else -1)
            for m, n, k in CONFIG_SHAPES:
                a = rng.uniform(-0.5, 0.5, (m, k))
                b = rng.uniform(-0.5, 0.5, (k, n))
                if complex_:
                    a = a + 1j * rng.uniform(-0.5, 0.5, (m, k))
                    b = b + 1j * rng.uniform(-0.5, 0.5, (k, n))
                c = ws.multiply(a, b)
                ref = O.gemm(a, b)
                assert np.all(np.abs(c - ref) <= 4 * bound(a, b, 0 * ref, 1, 0, k)), (cfg, m, n, k)
    finally:
        lib.tns_gemm_set_config(-1, -1)


@pytest.mark.parametrize("offset", [0, 1])
def test_device_operands_at_odd_offsets(ws, rng, offset):
    # Device operands that start 8 bytes past a 16-byte boundary, or have an
    # odd leading dimension, must take the element-wise loads.
    import torch

    import csharp_linearalgebra_amd as T

    lib = T.load()
    m, n, k = 256, 128, 64
    for ld_extra in (0, 1):
        lda, ldb, ldc = m + ld_extra, k + ld_extra, m
        a = rng.uniform(-0.5, 0.5, (m, k))
        b = rng.uniform(-0.5, 0.5, (k, n))
        abuf = torch.zeros(offset + lda * k, dtype=torch.float64, device="cuda")
        bbuf = torch.zeros(offset + ldb * n, dtype=torch.float64, device="cuda")
        cbuf = torch.zeros(ldc * n, dtype=torch.float64, device="cuda")
        abuf[offset:].view(k, lda)[:, :m] = torch.from_numpy(a.T.copy()).cuda()
        bbuf[offset:].view(n, ldb)[:, :k] = torch.from_numpy(b.T.copy()).cuda()
        torch.cuda.synchronize()
        rc = lib.tns_dgemm(ws.handle, m, n, k, 1.0, abuf.data_ptr() + 8 * offset, lda,
                           bbuf.data_ptr() + 8 * offset, ldb, 0.0, cbuf.data_ptr(), ldc)
        assert rc == 0
        ws.synchronize()
        c = cbuf.view(n, ldc).cpu().numpy().T
        ref = O.gemm(a, b)
        assert np.all(np.abs(c - ref) <= 2 * bound(a, b, 0 * ref, 1, 0, k)), (offset, ld_extra)
