"""Device norms, panel products, norm estimation, condition estimation, the
matrix exponential and its action, against the oracle and the 50-digit
golden values (NormEstimateTests.cs, MatrixExponentialTests.cs,
MatrixExponentialActionTests.cs, ComplexExponentialTests.cs).

Tolerances: Expm / Expmv relative 1-norm error within 10x the oracle's own
error against the extended-precision reference, with a
floor of 64 eps for cases where the oracle's error is zero to rounding."""
import os

import numpy as np
import pytest
import scipy.linalg as sl

import oracle as O

pytestmark = pytest.mark.gpu
EPS = np.finfo(float).eps
GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "dense_path_golden.npz"))


def rel1(x, ref):
    return np.abs(x - ref).sum(0).max() / np.abs(ref).sum(0).max()


@pytest.mark.parametrize("m,n", [(1, 1), (7, 3), (300, 200), (1000, 1000), (4097, 33)])
def test_norms(ws, rng, m, n):
    a = rng.uniform(-1, 1, (m, n))
    z = a + 1j * rng.uniform(-1, 1, (m, n))
    assert ws.one_norm(a) == pytest.approx(O.norm(a, "1"), rel=1e-13)
    assert ws.infinity_norm(a) == pytest.approx(O.norm(a, "inf"), rel=1e-13)
    assert ws.frobenius_norm(a) == pytest.approx(O.norm(a, "fro"), rel=1e-13)
    assert ws.one_norm(z) == pytest.approx(O.norm(z, "1"), rel=1e-13)
    assert ws.infinity_norm(z) == pytest.approx(O.norm(z, "inf"), rel=1e-13)


def test_norm_of_empty_and_nan(ws):
    assert ws.one_norm(np.zeros((0, 5))) == 0.0
    a = np.ones((4, 4))
    a[2, 1] = np.nan
    assert np.isnan(ws.one_norm(a)) and np.isnan(ws.infinity_norm(a))


# 3000 x 3 and 2049 x 8 take X through LDS in several stages, the last one
# partial (column chunks of 273 and 129 in stages of 64; row stages of 512 or
# 1024, the last of 2049 one row long).
@pytest.mark.parametrize("n,t", [(1, 1), (50, 2), (1000, 2), (1000, 16), (777, 5), (300, 40), (3000, 3), (2049, 8)])
def test_panel_products(ws, rng, n, t):
    a = rng.uniform(-1, 1, (n, n))
    x = rng.uniform(-1, 1, (n, t))
    for tr in (False, True):
        y = ws.multiply_panel(a, x, tr)
        ref = O.panel(a, x, tr)
        assert np.max(np.abs(y - ref)) <= (n + 2) * EPS * np.max(np.abs(a) @ np.abs(x) if not tr else np.abs(a.T) @ np.abs(x))
    z = a + 1j * rng.uniform(-1, 1, (n, n))
    xz = x + 1j * rng.uniform(-1, 1, (n, t))
    assert np.allclose(ws.multiply_panel(z, xz), z @ xz, rtol=0, atol=4 * (n + 2) * EPS * n)
    assert np.allclose(ws.multiply_panel(z, xz, True), z.conj().T @ xz, rtol=0, atol=4 * (n + 2) * EPS * n)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 16, 17, 40, 64, 500])
def test_normest_matches_the_oracle(ws, n):
    for trial in range(3):
        r = np.random.default_rng(n * 10 + trial)
        a = r.uniform(-0.5, 0.5, (n, n))
        truth = np.linalg.norm(a, 1)
        got = ws.estimate_one_norm(a)
        ref = O.normest_power(a)
        assert got.value == pytest.approx(ref[0], rel=1e-12)
        assert (got.iterations, got.products) == ref[1:]
        assert got.value <= truth * (1 + 1e-12)
        assert ws.estimate_one_norm(a, columns=n).value == pytest.approx(truth, rel=1e-12)
        for p in (2, 5):
            assert ws.estimate_one_norm(a, power=p).value == pytest.approx(O.normest_power(a, p)[0], rel=1e-10)
        z = a + 1j * r.uniform(-0.5, 0.5, (n, n))
        gz = ws.estimate_one_norm(z, power=3)
        rz = O.normest_power(z, 3)
        assert gz.value == pytest.approx(rz[0], rel=1e-10)


@pytest.mark.parametrize("n", [5, 40, 200, 1024])
def test_gecon_matches_the_oracle(ws, n):
    r = np.random.default_rng(n)
    a = r.uniform(-1, 1, (n, n))
    lu = ws.factor_lu(a)
    ref = O.getrf(a)
    assert lu.reciprocal_condition() == pytest.approx(O.gecon(ref), rel=1e-9)
    z = a + 1j * r.uniform(-1, 1, (n, n))
    zl = ws.factor_lu(z)
    assert zl.reciprocal_condition() == pytest.approx(O.gecon(O.getrf(z)), rel=1e-9)
    if n <= 200:
        exact = 1 / (np.linalg.norm(a, 1) * np.linalg.norm(np.linalg.inv(a), 1))
        assert exact * (1 - 1e-10) <= lu.reciprocal_condition() <= 3 * exact
    assert ws.factor_lu(np.eye(7)).reciprocal_condition() == pytest.approx(1.0)
    assert ws.factor_lu(sl.hilbert(12)).reciprocal_condition() < 1e-12


@pytest.mark.parametrize("name", ["e0", "e1", "e2", "e3", "e4", "e5", "e6"])
def test_expm_against_golden(ws, name):
    a, ref = GOLD[f"{name}_a"], GOLD[f"{name}_exp"]
    e, d = ws.expm(a)
    eo, do = O.expm(a)
    assert (d.degree, d.squarings) == do
    assert rel1(e, ref) <= max(10 * rel1(eo, ref), 64 * EPS)


@pytest.mark.parametrize("n,norm,cplx", [(8, 0.4, False), (17, 12.0, False), (32, 30.0, False), (64, 100.0, False),
                                         (50, 2.5, True), (120, 40.0, True), (300, 8.0, False), (300, 8.0, True),
                                         (513, 20.0, False), (257, 20.0, True)])
def test_expm_against_the_oracle(ws, n, norm, cplx):
    r = np.random.default_rng(n)
    a = r.standard_normal((n, n)) + (1j * r.standard_normal((n, n)) if cplx else 0)
    a *= norm / np.linalg.norm(a, 1)
    e, d = ws.expm(a)
    eo, do = O.expm(a)
    assert (d.degree, d.squarings) == do
    if n <= 64:
        ref = O.expm_taylor_ld(a)
        assert rel1(e, ref) <= max(10 * rel1(eo, ref), 64 * EPS)
    else:
        assert rel1(e, eo) <= 1e-12


def test_expm_degrees_and_closed_forms(ws):
    seen = set()
    for norm in (0.001, 0.1, 0.5, 1.5, 3.0, 50.0):
        a = np.random.default_rng(1).standard_normal((20, 20))
        a *= norm / np.linalg.norm(a, 1)
        seen.add(ws.expm(a)[1].degree)
    assert seen == {3, 5, 7, 9, 13}
    assert np.array_equal(ws.expm(np.zeros((6, 6)))[0], np.eye(6))
    th = 2.5
    e, _ = ws.expm(np.array([[0.0, -th], [th, 0.0]]))
    assert np.allclose(e, [[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]], atol=1e-15)
    assert ws.expm(np.array([[1.5 + 2j]]))[0][0, 0] == pytest.approx(np.exp(1.5 + 2j), rel=1e-15)
    assert ws.expm(np.zeros((0, 0)))[0].shape == (0, 0)


def test_config2_expm(ws):
    from csharp_linearalgebra_amd import workloads as W

    a, _ = W.expm_operator(1000, 50.0, 8, seed=0)
    e, d = ws.expm(a)
    eo, do = O.expm(a)
    assert (d.degree, d.squarings) == do
    assert rel1(e, eo) <= 1e-12
    # exp(-A) exp(A) = I to the conditioning of the pair.
    em, _ = ws.expm(-a)
    assert np.abs(e @ em - np.eye(1000)).max() <= 1e-9


@pytest.mark.parametrize("norm", [0.05, 2.0, 50.0])
def test_expm_phase_split(ws, norm):
    """The timed Expm returns the same exponential bit for bit, and its
    phases add up to at most its total (each phase is a span between two
    events on the stream, the total the span from the first to the last)."""
    import csharp_linearalgebra_amd as T

    r = np.random.default_rng(3)
    n = 200
    a = r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))
    a *= norm / np.linalg.norm(a, 1)
    e0, d0 = ws.expm(a)
    ph = T.ExpmPhaseTimings()
    e1, d1 = ws.expm(a, phases=ph)
    assert np.array_equal(e0, e1) and d0 == d1
    parts = [ph.powers_ms, ph.select_ms, ph.pade_ms, ph.lu_ms, ph.solve_ms, ph.squaring_ms]
    assert all(p >= 0.0 for p in parts) and ph.total_ms > 0.0
    assert sum(parts) <= ph.total_ms * (1 + 1e-6) + 1e-3
    assert ph.lu_ms > 0.0 and ph.solve_ms > 0.0 and ph.pade_ms > 0.0
    assert ph.calls == 1 and ph.per_call()["total"] == pytest.approx(ph.total_ms, abs=1e-3)


@pytest.mark.parametrize("name", ["e0", "e1", "e2", "e3", "e4", "e5", "e6"])
def test_expmv_against_golden(ws, name):
    a, b, ref = GOLD[f"{name}_a"], GOLD[f"{name}_b"], GOLD[f"{name}_expmv"]
    f, d = ws.expmv(a, b, 0.5)
    fo, do = O.expmv(a, b, 0.5)
    assert (d.degree, d.scaling, d.applications, d.early_exits) == do
    err, err_o = np.abs(f - ref).max(), np.abs(fo - ref).max()
    assert err <= max(10 * err_o, 64 * EPS * np.abs(ref).max())


def test_config2_expmv(ws):
    from csharp_linearalgebra_amd import workloads as W

    a, b = W.expm_operator(1000, 50.0, 8, seed=0)
    f, d = ws.expmv(a, b, 0.5)
    fo, do = O.expmv(a, b, 0.5)
    assert (d.degree, d.scaling) == do[:2]
    assert np.abs(f - fo).max() <= 1e-12 * np.abs(fo).max()


def test_expmv_properties(ws):
    r = np.random.default_rng(11)
    n = 40
    a = r.standard_normal((n, n)) * 0.3
    b = r.standard_normal((n, 2))
    f, _ = ws.expmv(a, b, 1.0)
    half, _ = ws.expmv(a, b, 0.5)
    twice, _ = ws.expmv(a, half, 0.5)
    assert np.allclose(f, twice, rtol=1e-12, atol=1e-12)
    assert np.allclose(ws.expmv(np.zeros((n, n)), b, 2.0)[0], b)
    h = r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))
    skew = h - h.conj().T
    fz, _ = ws.expmv(skew, b.astype(complex), 1.0)
    assert np.allclose(np.linalg.norm(fz, axis=0), np.linalg.norm(b, axis=0), rtol=1e-12)
    fneg, _ = ws.expmv(a, b, -0.75)
    assert np.allclose(fneg, sl.expm(-0.75 * a) @ b, rtol=1e-12, atol=1e-12)


def test_matrix_free_action(ws):
    import csharp_linearalgebra_amd as T

    r = np.random.default_rng(3)
    n = 60
    a = r.standard_normal((n, n)) * 0.2
    b = r.standard_normal((n, 3))

    class Dense(T.ILinearOperator):
        order = n

        def apply(self, x):
            return a @ x

    f, d = ws.expmv_operator(Dense(), b, 1.0, np.linalg.norm(a, 1))
    fo, do = O.expmv_op(n, lambda x: a @ x, b, 1.0, np.linalg.norm(a, 1))
    assert (d.degree, d.scaling) == do[:2]
    assert np.allclose(f, fo, rtol=1e-12, atol=1e-13)
    with pytest.raises(T.ArgumentOutOfRangeError):
        ws.expmv_operator(Dense(), b, 1.0, float("nan"))

    class Broken(T.ILinearOperator):
        order = n

        def apply(self, x):
            raise KeyError("boom")

    with pytest.raises(T.OperatorError):
        ws.expmv_operator(Broken(), b, 1.0, 1.0)
