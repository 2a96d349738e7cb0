"""Pins the oracle (the CPU restatement of Tensile's algorithms) before any
device result is compared with it: against LAPACK and 50-digit mpmath
fixtures (tests/golden), closed forms, and the invariants the reference's
own tests assert. CPU only."""
import os

import numpy as np
import pytest
import scipy.linalg as sl

import oracle as O

EPS = np.finfo(float).eps
GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "dense_path_golden.npz"))


def rng(seed=0):
    return np.random.default_rng(seed)


# ---- GEMM -------------------------------------------------------------------------
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (17, 9, 5), (65, 63, 66), (100, 37, 129), (300, 257, 513)])
def test_gemm_matches_the_textbook_loop(m, n, k):
    r = rng(m * n + k)
    a, b = r.uniform(-0.5, 0.5, (m, k)), r.uniform(-0.5, 0.5, (k, n))
    c = O.gemm(a, b)
    exact = (a.astype(np.longdouble) @ b.astype(np.longdouble)).astype(float)
    assert np.all(np.abs(c - exact) <= (k + 2) * EPS * (np.abs(a) @ np.abs(b)))
    assert np.max(np.abs(O.gemm(a, b, reference=True) - exact)) <= (k + 2) * EPS * np.max(np.abs(a) @ np.abs(b))


@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-1.0, 1.0), (0.0, 1.0), (0.0, 0.0), (2.5, -0.75)])
def test_gemm_scalar_semantics(alpha, beta):
    r = rng(3)
    a, b, c = r.uniform(-1, 1, (37, 41)), r.uniform(-1, 1, (41, 29)), r.uniform(-1, 1, (37, 29))
    got = O.gemm(a, b, c, alpha, beta)
    assert np.allclose(got, beta * c + alpha * (a @ b), rtol=0, atol=1e-13)


def test_gemm_beta_zero_overwrites_nan():
    a, b = np.ones((5, 3)), np.ones((3, 4))
    c = np.full((5, 4), np.nan)
    assert np.all(O.gemm(a, b, c, 1.0, 0.0) == 3.0)
    assert np.all(O.gemm(a[:, :0], b[:0], c, 1.0, 0.0) == 0.0)


def test_zgemm_4m_is_the_complex_product_and_keeps_small_imaginary_parts():
    r = rng(5)
    a = r.uniform(-1, 1, (40, 30)) + 1j * r.uniform(-1, 1, (40, 30))
    b = r.uniform(-1, 1, (30, 20)) + 1j * r.uniform(-1, 1, (30, 20))
    assert np.max(np.abs(O.gemm(a, b) - a @ b)) <= 64 * EPS
    n = 64
    for rho in (1e6, 1e9, 1e12):
        a = r.uniform(0.5, 1, (n, n)) + 1j * r.uniform(-1, 1, (n, n)) / rho
        b = r.uniform(0.5, 1, (n, n)) + 1j * r.uniform(-1, 1, (n, n)) / rho
        exact = a.real @ b.imag + a.imag @ b.real
        scale = np.abs(a.real) @ np.abs(b.imag) + np.abs(a.imag) @ np.abs(b.real)
        assert np.max(np.abs(O.gemm(a, b).imag - exact) / scale) <= 4 * (n + 2) * EPS
        assert np.max(np.abs(O.zgemm_3m(a, b).imag - exact) / scale) > 4 * (n + 2) * EPS


@pytest.mark.parametrize("alpha,beta", [(1, 0), (2.5, 0), (0.5 - 1.5j, 0), (1, 1), (0.3 + 0.7j, 0.9 - 0.1j)])
def test_zgemm_combine_cases(alpha, beta):
    r = rng(6)
    a = r.uniform(-1, 1, (9, 7)) + 1j * r.uniform(-1, 1, (9, 7))
    b = r.uniform(-1, 1, (7, 5)) + 1j * r.uniform(-1, 1, (7, 5))
    c = r.uniform(-1, 1, (9, 5)) + 1j * r.uniform(-1, 1, (9, 5))
    assert np.allclose(O.gemm(a, b, c, alpha, beta), beta * c + alpha * (a @ b), rtol=0, atol=1e-14)


# ---- .NET arithmetic restatements ------------------------------------------------
def test_complex_modulus_and_division():
    for z in [3 + 4j, -1e300 + 1e300j, 1e-300 - 2e-300j, 0 + 0j, 5 - 0j, np.inf + 1j]:
        assert O.cabs(z) == pytest.approx(abs(z), rel=2 * EPS)
    for x, y in [(1 + 2j, 3 - 4j), (1 + 0j, 2e-300 + 1e-300j), (-7 + 1j, 0 + 3j)]:
        q = O.cdiv(x, y)
        assert abs(q - x / y) <= 4 * EPS * abs(x / y)


def test_random_is_deterministic_and_in_range():
    r1, r2 = O.Random(20260915), O.Random(20260915)
    s1 = [r1.next(2) for _ in range(2000)]
    s2 = [r2.next(2) for _ in range(2000)]
    assert s1 == s2 and set(s1) == {0, 1}
    assert 0.45 < np.mean(s1) < 0.55
    r3 = O.Random(7)
    xs = [r3.sample() for _ in range(5000)]
    assert 0.0 <= min(xs) and max(xs) < 1.0 and abs(np.mean(xs) - 0.5) < 0.02
    assert [O.Random(1).next(1000) for _ in range(3)] == [O.Random(1).next(1000)] * 3


@pytest.mark.parametrize("n", [2, 3, 16, 17, 40, 1000, 5000])
def test_introsort_sorts_and_permutes(n):
    r = rng(n)
    keys = r.integers(-20, 20, n).astype(float)  # many ties
    keys[r.integers(0, n, max(1, n // 50))] = np.nan
    k, v = O.sort_keys_items(keys, np.arange(n))
    nan = np.isnan(k)
    m = int(nan.sum())
    assert np.all(nan[:m]) and not np.any(nan[m:])
    assert np.all(np.diff(k[m:]) >= 0)
    assert sorted(v.tolist()) == list(range(n))
    assert np.array_equal(np.where(np.isnan(keys[v]), 0, keys[v]), np.where(np.isnan(k), 0, k))
    distinct = r.permutation(n).astype(float)
    k2, v2 = O.sort_keys_items(distinct, np.arange(n))
    assert np.array_equal(v2, np.argsort(distinct))


# ---- LU ---------------------------------------------------------------------------------
@pytest.mark.parametrize("idx", range(5))
def test_lu_pivots_and_factors_match_lapack(idx):
    for kind in ("d", "z"):
        a = GOLD[f"{kind}lu{idx}_a"]
        for nb in (0, 4, 8, 64):
            lu = O.getrf(a, nb)
            assert np.array_equal(lu.pivots, GOLD[f"{kind}lu{idx}_piv"][: min(a.shape)])
            assert np.max(np.abs(lu.factors - GOLD[f"{kind}lu{idx}_lu"])) <= 1e-12


def test_lu_kats():
    assert np.prod(np.diag(O.getrf(np.array([[1.0, 2.0], [3.0, 4.0]])).factors)) * -1 == pytest.approx(-2.0)
    lu = O.getrf(np.array([[3.0, 1.0], [2 + 2j, 1.0]]))
    assert lu.pivots[0] == 1  # |Re|+|Im|: 4 beats 3 though |3| > |2+2i|
    lu = O.getrf(np.array([[1 + 1j, 1], [-2.0 + 0j, 2], [0.5 - 1.5j, 3]]))
    assert lu.pivots[0] == 0  # ties to the lowest index
    lu = O.getrf(np.array([[1.0 + 0j, 0], [np.nan, 1], [2j, 0]]))
    assert lu.pivots[0] == 2  # a NaN never wins
    z = O.getrf(np.array([[0, 2j], [3, 0]]))
    det = np.prod(np.diag(z.factors)) * (-1) ** int(np.sum(z.pivots != np.arange(2)))
    assert abs(det - (-6j)) < 1e-15


def test_singular_column_is_the_first_zero_pivot():
    r = rng(12)
    n = 16
    a = r.uniform(-0.5, 0.5, (n, n)) + n * np.eye(n)
    a[:, 5] = 0.0
    assert O.getrf(a, 4).singular_column == 5
    assert O.gecon(O.getrf(a, 4)) == 0.0


@pytest.mark.parametrize("n,nrhs", [(1, 1), (50, 3), (200, 17)])
def test_solves_match_lapack(n, nrhs):
    r = rng(n)
    for cplx in (False, True):
        a = r.uniform(-1, 1, (n, n)) + (1j * r.uniform(-1, 1, (n, n)) if cplx else 0)
        b = r.uniform(-1, 1, (n, nrhs)) + (1j * r.uniform(-1, 1, (n, nrhs)) if cplx else 0)
        lu = O.getrf(a)
        assert np.allclose(O.getrs(lu, b), np.linalg.solve(a, b), rtol=1e-10, atol=1e-10)
        assert np.allclose(O.getrs(lu, b, True), np.linalg.solve(a.conj().T, b), rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("kind", range(6))
def test_triangular_solves_match_scipy(kind):
    r = rng(kind)
    n = 60
    a = r.uniform(-1, 1, (n, n)) * 0.1 + n * np.eye(n)
    b = r.uniform(-1, 1, (n, 5))
    lower, trans, unit = {0: (True, 0, True), 1: (False, 0, False), 2: (True, 0, False), 3: (False, 1, False),
                          4: (True, 1, False), 5: (True, 1, True)}[kind]
    ref = sl.solve_triangular(a, b, lower=lower, trans=trans, unit_diagonal=unit)
    assert np.allclose(O.trsv(kind, a, b), ref, rtol=1e-13, atol=1e-14)


# ---- norms, panels, estimator -------------------------------------------------------------
def test_norms_and_panels():
    r = rng(8)
    a = r.uniform(-1, 1, (30, 20))
    z = a + 1j * r.uniform(-1, 1, (30, 20))
    assert O.norm(a, "1") == pytest.approx(np.linalg.norm(a, 1), rel=1e-15)
    assert O.norm(a, "inf") == pytest.approx(np.linalg.norm(a, np.inf), rel=1e-15)
    assert O.norm(a, "fro") == pytest.approx(np.linalg.norm(a, "fro"), rel=1e-15)
    assert O.norm(z, "1") == pytest.approx(np.linalg.norm(z, 1), rel=1e-14)
    assert O.norm(z, "inf") == pytest.approx(np.linalg.norm(z, np.inf), rel=1e-14)
    s = r.uniform(-1, 1, (25, 25))
    x = r.uniform(-1, 1, (25, 3))
    assert np.allclose(O.panel(s, x), s @ x, atol=1e-14)
    assert np.allclose(O.panel(s, x, True), s.T @ x, atol=1e-14)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 16, 17, 40, 64])
def test_normest_invariants(n):
    # NormEstimateTests.cs:31-100: t = n exact, A >= 0 exact, never above.
    for trial in range(3):
        r = rng(n * 10 + trial)
        a = r.uniform(-0.5, 0.5, (n, n))
        truth = np.linalg.norm(a, 1)
        assert abs(O.normest_power(a, 1, columns=n)[0] - truth) <= 1e-12 * truth
        nonneg = r.uniform(0, 1, (n, n))
        assert abs(O.normest_power(nonneg)[0] - np.linalg.norm(nonneg, 1)) <= 1e-12 * np.linalg.norm(nonneg, 1)
        for t in (1, 2, 4):
            assert O.normest_power(a, 1, columns=t)[0] <= truth * (1 + 1e-12)
    a = rng(99).uniform(-0.5, 0.5, (64, 64))
    assert O.normest_power(a) == O.normest_power(a)


@pytest.mark.parametrize("n,power", [(8, 2), (20, 3), (33, 5)])
def test_power_operator_matches_the_explicit_power(n, power):
    a = rng(n).uniform(-0.5, 0.5, (n, n))
    ap = np.linalg.matrix_power(a, power)
    assert abs(O.normest_power(a, power, columns=n)[0] - np.linalg.norm(ap, 1)) <= 1e-10 * np.linalg.norm(ap, 1)
    z = a + 1j * rng(n + 1).uniform(-0.5, 0.5, (n, n))
    zp = np.linalg.matrix_power(z, power)
    assert abs(O.normest_power(z, power, columns=n)[0] - np.linalg.norm(zp, 1)) <= 1e-10 * np.linalg.norm(zp, 1)


def test_complex_estimate_equals_real_on_real_input_and_is_phase_invariant():
    a = rng(4).uniform(-1, 1, (30, 30))
    truth = np.linalg.norm(a, 1)
    assert O.normest_power(a.astype(complex))[0] <= truth * (1 + 1e-12)
    ph = np.exp(1j * rng(5).uniform(0, 2 * np.pi, 30))
    z = a * ph[:, None]
    assert O.normest_power(z, 1, columns=30)[0] == pytest.approx(truth, rel=1e-12)


def test_gecon_against_exact_and_lapack():
    for idx in range(2):
        a = GOLD[f"dgecon{idx}_a"]
        rc = O.gecon(O.getrf(a))
        exact = GOLD[f"dgecon{idx}_exact"]
        assert exact * (1 - 1e-10) <= rc <= 3 * exact
        assert 0.3 < rc / GOLD[f"dgecon{idx}_rc"] < 3.0
    assert O.gecon(O.getrf(np.eye(7))) == pytest.approx(1.0)
    h = sl.hilbert(12)
    assert O.gecon(O.getrf(h)) < 1e-12
    for n in (5, 20, 40):
        a = rng(n).uniform(-1, 1, (n, n))
        exact = 1 / (np.linalg.norm(a, 1) * np.linalg.norm(np.linalg.inv(a), 1))
        rc = O.gecon(O.getrf(a))
        assert exact * (1 - 1e-10) <= rc <= 3 * exact
        n_full = O.normest_op(n, lambda x: np.linalg.solve(a, x), lambda x: np.linalg.solve(a.T, x), columns=n)[0]
        assert n_full == pytest.approx(np.linalg.norm(np.linalg.inv(a), 1), rel=1e-10)


# ---- matrix exponential ----------------------------------------------------------------------
def test_pade_coefficients_match_their_closed_form():
    from fractions import Fraction
    from math import factorial

    # b_j = (2m-j)! m! / ((2m)! j! (m-j)!), normalised so b_m = 1, in exact
    # arithmetic (MatrixExponentialTests.cs:51-68); every one is an integer
    # exactly representable in a double (MatrixExponentialTests.cs:82-100).
    tables = {3: [120, 60, 12, 1], 5: [30240, 15120, 3360, 420, 30, 1],
              7: [17297280, 8648640, 1995840, 277200, 25200, 1512, 56, 1],
              9: [17643225600, 8821612800, 2075673600, 302702400, 30270240, 2162160, 110880, 3960, 90, 1],
              13: [64764752532480000, 32382376266240000, 7771770303897600, 1187353796428800, 129060195264000,
                   10559470521600, 670442572800, 33522128640, 1323241920, 40840800, 960960, 16380, 182, 1]}
    for m, tab in tables.items():
        b = [Fraction(factorial(2 * m - j) * factorial(m), factorial(2 * m) * factorial(j) * factorial(m - j))
             for j in range(m + 1)]
        assert [x / b[m] for x in b] == [Fraction(t) for t in tab]
        assert all(float(t) == t for t in tab)
    # The oracle carries these tables: a degree-3 Pade of a small scalar
    # agrees with exp to the approximant's own error.
    x = 0.01
    p = sum(c * x ** j for j, c in enumerate(tables[3]))
    q = sum(c * (-x) ** j for j, c in enumerate(tables[3]))
    assert O.expm(np.array([[x, 0.0], [0.0, x]]) * 1.0)[0][0, 0] == pytest.approx(p / q, rel=1e-15)


@pytest.mark.parametrize("name", ["e0", "e1", "e2", "e3", "e4", "e5", "e6"])
def test_expm_against_50_digit_reference(name):
    a, ref = GOLD[f"{name}_a"], GOLD[f"{name}_exp"]
    e, _ = O.expm(a)
    rel = np.abs(e - ref).sum(0).max() / np.abs(ref).sum(0).max()
    assert rel <= 1e-13, rel
    et = O.expm_taylor_ld(a)
    assert np.abs(et - ref).sum(0).max() / np.abs(ref).sum(0).max() <= 1e-15


def test_expm_closed_forms():
    n = 6
    assert np.array_equal(O.expm(np.zeros((n, n)))[0], np.eye(n))
    for scale in (1e-4, 0.5, 3.0, 20.0):
        d = np.diag(np.array([0.0, 1.0, -1.0, 0.25, -2.0]) * scale)
        want = np.diag(np.exp(np.diag(d)))
        assert np.all(np.abs(O.expm(d)[0] - want) <= 1e-11 * np.maximum(np.abs(want), 1.0))
    th = 2.5
    rot = np.array([[0.0, -th], [th, 0.0]])
    assert np.allclose(O.expm(rot)[0], [[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]], atol=1e-15)
    lam, k = 0.7, 5
    j = lam * np.eye(k) + np.eye(k, k=1)
    from math import factorial

    want = np.exp(lam) * sum(np.linalg.matrix_power(np.eye(k, k=1), p) / factorial(p) for p in range(k))
    assert np.allclose(O.expm(j)[0], want, rtol=1e-13, atol=0)
    nil = np.triu(rng(2).uniform(-1, 1, (8, 8)), 1) * 5
    want = sum(np.linalg.matrix_power(nil, p) / factorial(p) for p in range(8))
    assert np.allclose(O.expm(nil)[0], want, rtol=1e-12, atol=1e-12)
    z = np.array([[1.5 + 2j]])
    assert O.expm(z)[0][0, 0] == pytest.approx(np.exp(1.5 + 2j), rel=1e-15)


def test_every_pade_degree_is_reached():
    seen = set()
    for norm in (0.001, 0.1, 0.5, 1.5, 3.0, 50.0):
        a = rng(1).standard_normal((20, 20))
        a *= norm / np.linalg.norm(a, 1)
        seen.add(O.expm(a)[1][0])
    assert seen == {3, 5, 7, 9, 13}


@pytest.mark.parametrize("norm", [0.2, 3.0, 30.0])
def test_expm_identities(norm):
    a = rng(3).standard_normal((24, 24))
    a *= norm / np.linalg.norm(a, 1)
    e, _ = O.expm(a)
    em, _ = O.expm(-a)
    assert np.abs(e @ em - np.eye(24)).max() <= 1e-12 * max(1, np.abs(e).max() * np.abs(em).max())
    assert np.allclose(O.expm(a.T)[0], e.T, rtol=1e-12, atol=1e-14 * np.abs(e).max())
    assert np.allclose(sl.expm(a), e, rtol=1e-12, atol=1e-13 * np.abs(e).max())


# ---- exponential action ------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["e0", "e1", "e2", "e3", "e4", "e5", "e6"])
def test_expmv_against_50_digit_reference(name):
    a, b, ref = GOLD[f"{name}_a"], GOLD[f"{name}_b"], GOLD[f"{name}_expmv"]
    f, _ = O.expmv(a, b, 0.5)
    assert np.abs(f - ref).max() <= 1e-12 * max(1, np.abs(ref).max())


def test_expmv_properties():
    r = rng(11)
    n = 40
    a = r.standard_normal((n, n)) * 0.3
    b = r.standard_normal((n, 2))
    f, _ = O.expmv(a, b, 1.0)
    half, _ = O.expmv(a, b, 0.5)
    twice, _ = O.expmv(a, half, 0.5)
    assert np.allclose(f, twice, rtol=1e-12, atol=1e-12)
    assert np.allclose(O.expmv(np.zeros((n, n)), b, 2.0)[0], b)
    d = np.diag(r.uniform(-2, 2, n))
    assert np.allclose(O.expmv(d, b, 0.7)[0], np.exp(0.7 * np.diag(d))[:, None] * b, rtol=1e-13)
    h = r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))
    skew = h - h.conj().T
    bz = b.astype(complex)
    fz, _ = O.expmv(skew, bz, 1.0)
    assert np.allclose(np.linalg.norm(fz, axis=0), np.linalg.norm(bz, axis=0), rtol=1e-12)
    fop, _ = O.expmv_op(n, lambda x: a @ x, b, 1.0, np.linalg.norm(a, 1))
    assert np.allclose(fop, f, rtol=1e-12, atol=1e-12)


def test_expmv_parameter_selection_matches_scipy_on_the_norm_bound_branch():
    # Condition 3.13 holds for small ||tA||_1, where the (m, s) choice is the
    # plain minimisation over the theta table (Al-Mohy & Higham 2011,
    # Algorithm 3.2), the same one scipy's expm_multiply makes.
    from scipy.sparse.linalg import _expm_multiply as em

    r = rng(12)
    for norm in (0.01, 0.3, 1.0, 4.0, 9.0):
        n = 30
        a = r.standard_normal((n, n))
        a -= np.trace(a) / n * np.eye(n)
        a *= norm / np.linalg.norm(a, 1)
        b = r.standard_normal((n, 1))
        _, (m, s, _, _) = O.expmv(a, b, 1.0)
        info = em.LazyOperatorNormInfo(a, A_1_norm=np.linalg.norm(a, 1), ell=2)
        m_ref, s_ref = em._fragment_3_1(info, 1, 2 ** -53, ell=2)
        assert (m, s) == (m_ref, s_ref), (norm, (m, s), (m_ref, s_ref))
