"""Gradients through F = exp(tA) B (the adjoint of Expmv): checked against
central finite differences of the device forward (h = 1e-6) and against the exact derivative from the 2n x 2n block exponential
(Frechet derivative: exp([[X, E], [0, X]]) has L(X, E) in its corner),
computed by the oracle's Expm (MatrixExponential.cs:181-270 restated).
Loss L = ||C F - Y||_F^2, G = dL/dF = 2 C^H (C F - Y)."""
import numpy as np
import pytest

import oracle as O
from csharp_linearalgebra_amd import workloads as W

pytestmark = pytest.mark.gpu


def loss(ws, a, b, c, y, t):
    f, _ = ws.expmv(a, b, t)
    r = c @ f - y
    return float(np.sum(np.abs(r) ** 2)), f


def references(a, b, g, t):
    n = a.shape[0]
    ah = a.conj().T
    blk = np.zeros((2 * n, 2 * n), dtype=np.result_type(a, g))
    blk[:n, :n] = t * ah
    blk[n:, n:] = t * ah
    blk[:n, n:] = g @ b.conj().T
    e, _ = O.expm(blk)
    ga = t * e[:n, n:]
    gb = e[:n, :n] @ g
    return ga, gb


@pytest.mark.parametrize("complex_", [False, True])
@pytest.mark.parametrize("n", [48, 130])
def test_gradient_matches_the_block_exponential(ws, complex_, n):
    a, b, c, y = W.adjoint_problem(n, complex_, target=20.0, seed=n)
    t = 1.0
    _, f = loss(ws, a, b, c, y, t)
    g = 2.0 * c.conj().T @ (c @ f - y)
    ga, gb = ws.expmv_backward(a, b, t, g)
    ra, rb = references(a, b, g, t)
    assert np.abs(ga - ra).max() <= 1e-10 * np.abs(ra).max(), np.abs(ga - ra).max() / np.abs(ra).max()
    assert np.abs(gb - rb).max() <= 1e-10 * np.abs(rb).max()


@pytest.mark.parametrize("complex_", [False, True])
def test_gradient_matches_central_differences(ws, complex_):
    n = 40
    a, b, c, y = W.adjoint_problem(n, complex_, target=20.0, seed=3)
    t = 1.0
    l0, f = loss(ws, a, b, c, y, t)
    g = 2.0 * c.conj().T @ (c @ f - y)
    ga, gb = ws.expmv_backward(a, b, t, g)
    rng = np.random.default_rng(5)
    h = 1e-6
    for _ in range(3):
        ea = rng.standard_normal(a.shape)
        eb = rng.standard_normal(b.shape)
        if complex_:
            ea = ea + 1j * rng.standard_normal(a.shape)
            eb = eb + 1j * rng.standard_normal(b.shape)
        lp, _ = loss(ws, a + h * ea, b + h * eb, c, y, t)
        lm, _ = loss(ws, a - h * ea, b - h * eb, c, y, t)
        fd = (lp - lm) / (2 * h)
        an = float(np.real(np.vdot(ga, ea) + np.vdot(gb, eb)))
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (fd, an)


def test_autograd_style_function(ws):
    import csharp_linearalgebra_amd as T

    a, b, c, y = W.adjoint_problem(32, True, seed=8)
    fn = T.ExpmvFunction(ws)
    f = fn.forward(a, b, 0.5)
    g = 2.0 * c.conj().T @ (c @ f - y)
    ga, gb = fn.backward(g)
    ga2, gb2 = ws.expmv_backward(a, b, 0.5, g)
    assert np.array_equal(ga, ga2) and np.array_equal(gb, gb2)
    with pytest.raises(T.InvalidOperationError):
        T.ExpmvFunction(ws).backward(g)


def test_zero_time_and_zero_matrix(ws):
    n = 12
    rng = np.random.default_rng(1)
    b = rng.standard_normal((n, 2))
    g = rng.standard_normal((n, 2))
    a = rng.standard_normal((n, n))
    ga, gb = ws.expmv_backward(a, b, 0.0, g)
    assert np.allclose(gb, g) and np.abs(ga).max() == 0.0
    ga, gb = ws.expmv_backward(np.zeros((n, n)), b, 1.0, g)
    assert np.allclose(gb, g)
    assert np.allclose(ga, g @ b.T, atol=1e-14)
