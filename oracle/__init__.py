"""ctypes view of the oracle: the CPU restatement of Tensile's dense hot path.

Test infrastructure. Only the test suite, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` import this package, and only as the checker
(or, for the baseline, as the thing timed on the host). The shipped library
never imports it.

All matrices are numpy arrays in Fortran (column-major) order, float64 or
complex128, matching Tensile's column-major ``MatrixView<T>``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")


class NormEst(C.Structure):
    _fields_ = [("value", C.c_double), ("iterations", C.c_int), ("products", C.c_int)]


class ExpmDiag(C.Structure):
    _fields_ = [("degree", C.c_int), ("squarings", C.c_int)]


class ExpmvDiag(C.Structure):
    _fields_ = [("degree", C.c_int), ("scaling", C.c_int), ("applications", C.c_int), ("early_exits", C.c_int)]


class Cplx(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


APPLY_D = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double))
APPLY_Z = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p)


def build() -> str:
    """Compile the oracle with its Makefile; returns the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    d, i, p = C.c_double, C.c_int, C.c_void_p
    L.or_dgemm.argtypes = [i, i, i, d, p, i, p, i, d, p, i]
    L.or_dgemm_ref.argtypes = [i, i, i, d, p, i, p, i, d, p, i]
    L.or_zgemm.argtypes = [i, i, i, Cplx, p, i, p, i, Cplx, p, i]
    L.or_zgemm_3m.argtypes = [i, i, i, p, i, p, i, p, i]
    L.or_set_threads.argtypes = [i]
    L.or_get_threads.restype = i
    for nm in ("or_dnorm1", "or_dnorminf", "or_dnormfro", "or_znorm1", "or_znorminf"):
        getattr(L, nm).argtypes = [i, i, p, i]
        getattr(L, nm).restype = d
    L.or_dpanel.argtypes = [i, i, i, p, i, p, i, p, i]
    L.or_dgetrf.argtypes = [i, i, p, i, i, p, p, p]
    L.or_dgetrf.restype = i
    L.or_zgetrf.argtypes = [i, i, p, i, i, p, p, p]
    L.or_zgetrf.restype = i
    L.or_dgetrs.argtypes = [i, i, i, p, i, p, p, i]
    L.or_zgetrs.argtypes = [i, i, i, p, i, p, p, i]
    L.or_dtrsv.argtypes = [i, i, i, p, i, p, i]
    L.or_dnormest.argtypes = [i, APPLY_D, APPLY_D, p, i, i, i]
    L.or_dnormest.restype = NormEst
    L.or_znormest.argtypes = [i, APPLY_Z, APPLY_Z, p, i, i, i]
    L.or_znormest.restype = NormEst
    L.or_dnormest_power.argtypes = [i, p, i, i, i, i, i]
    L.or_dnormest_power.restype = NormEst
    L.or_znormest_power.argtypes = [i, p, i, i, i, i, i]
    L.or_znormest_power.restype = NormEst
    L.or_dgecon.argtypes = [i, p, i, p, i, d, i]
    L.or_dgecon.restype = d
    L.or_zgecon.argtypes = [i, p, i, p, i, d, i]
    L.or_zgecon.restype = d
    L.or_random_init.argtypes = [p, i]
    L.or_random_sample.argtypes = [p]
    L.or_random_sample.restype = d
    L.or_random_next.argtypes = [p, i]
    L.or_random_next.restype = i
    L.or_sort_keys_items.argtypes = [i, p, p]
    L.or_dexpm.argtypes = [i, p, i, p, i, C.POINTER(ExpmDiag)]
    L.or_dexpm.restype = i
    L.or_zexpm.argtypes = [i, p, i, p, i, C.POINTER(ExpmDiag)]
    L.or_zexpm.restype = i
    L.or_dexpmv.argtypes = [i, i, p, i, p, i, d, p, i, C.POINTER(ExpmvDiag)]
    L.or_dexpmv.restype = i
    L.or_zexpmv.argtypes = [i, i, p, i, p, i, d, p, i, C.POINTER(ExpmvDiag)]
    L.or_zexpmv.restype = i
    L.or_dexpmv_op.argtypes = [i, APPLY_D, p, i, p, i, d, d, p, i, C.POINTER(ExpmvDiag)]
    L.or_dexpmv_op.restype = i
    L.or_zexpmv_op.argtypes = [i, APPLY_Z, p, i, p, i, d, d, p, i, C.POINTER(ExpmvDiag)]
    L.or_zexpmv_op.restype = i
    L.or_dexpm_taylor_ld.argtypes = [i, p, i, p, i]
    L.or_zexpm_taylor_ld.argtypes = [i, p, i, p, i]
    L.or_zexpmv_taylor_ld.argtypes = [i, p, i, d, i, p, i, p, i]
    L.or_cabs.argtypes = [Cplx]
    L.or_cabs.restype = d
    L.or_cdiv.argtypes = [Cplx, Cplx]
    L.or_cdiv.restype = Cplx


def _f(a, dtype):
    """A Fortran-ordered copy (or view) of a, of the given dtype."""
    return np.asfortranarray(a, dtype=dtype)


def _ld(a):
    return max(1, a.strides[1] // a.itemsize) if a.ndim == 2 and a.shape[1] > 1 else max(1, a.shape[0])


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _cplx(z):
    z = complex(z)
    return Cplx(z.real, z.imag)


# ---- GEMM ------------------------------------------------------------------
def gemm(a, b, c=None, alpha=1.0, beta=0.0, reference=False):
    """C := beta*C + alpha*A*B; returns a new array (the input C is not modified)."""
    cplx = np.iscomplexobj(a) or np.iscomplexobj(b) or (c is not None and np.iscomplexobj(c)) \
        or isinstance(alpha, complex) or isinstance(beta, complex)
    dt = np.complex128 if cplx else np.float64
    a = _f(a, dt); b = _f(b, dt)
    m, k = a.shape
    k2, n = b.shape
    assert k == k2
    out = np.zeros((m, n), dtype=dt, order="F") if c is None else np.array(c, dtype=dt, order="F", copy=True)
    L = lib()
    if cplx:
        L.or_zgemm(m, n, k, _cplx(alpha), _ptr(a), max(1, m), _ptr(b), max(1, k), _cplx(beta), _ptr(out), max(1, m))
    else:
        fn = L.or_dgemm_ref if reference else L.or_dgemm
        fn(m, n, k, float(alpha), _ptr(a), max(1, m), _ptr(b), max(1, k), float(beta), _ptr(out), max(1, m))
    return out


def zgemm_3m(a, b):
    a = _f(a, np.complex128); b = _f(b, np.complex128)
    m, k = a.shape
    n = b.shape[1]
    out = np.zeros((m, n), dtype=np.complex128, order="F")
    lib().or_zgemm_3m(m, n, k, _ptr(a), max(1, m), _ptr(b), max(1, k), _ptr(out), max(1, m))
    return out


def set_threads(t: int):
    lib().or_set_threads(int(t))


def threads() -> int:
    return lib().or_get_threads()


# ---- norms -----------------------------------------------------------------
def norm(a, kind="1"):
    L = lib()
    if np.iscomplexobj(a):
        a = _f(a, np.complex128)
        fn = {"1": L.or_znorm1, "inf": L.or_znorminf}[kind]
    else:
        a = _f(a, np.float64)
        fn = {"1": L.or_dnorm1, "inf": L.or_dnorminf, "fro": L.or_dnormfro}[kind]
    m, n = a.shape
    return fn(m, n, _ptr(a), max(1, m))


def panel(a, x, trans=False):
    a = _f(a, np.float64); x = _f(x, np.float64)
    n = a.shape[0]
    t = x.shape[1]
    y = np.zeros((n, t), order="F")
    lib().or_dpanel(int(trans), n, t, _ptr(a), max(1, n), _ptr(x), max(1, n), _ptr(y), max(1, n))
    return y


# ---- LU ----------------------------------------------------------------------
class LU:
    def __init__(self, factors, pivots, singular, smallest, largest, anorm):
        self.factors = factors
        self.pivots = pivots
        self.singular_column = singular
        self.smallest = smallest
        self.largest = largest
        self.anorm = anorm

    @property
    def is_singular(self):
        return self.singular_column >= 0


def getrf(a, nb=0):
    cplx = np.iscomplexobj(a)
    dt = np.complex128 if cplx else np.float64
    f = np.array(a, dtype=dt, order="F", copy=True)
    m, n = f.shape
    piv = np.zeros(max(1, min(m, n)), dtype=np.int32)
    s = C.c_double()
    l = C.c_double()
    anorm = norm(f, "1") if m and n else 0.0
    fn = lib().or_zgetrf if cplx else lib().or_dgetrf
    sing = fn(m, n, _ptr(f), max(1, m), int(nb), _ptr(piv), C.byref(s), C.byref(l))
    return LU(f, piv[: min(m, n)], sing, s.value, l.value, anorm)


def getrs(lu: LU, b, trans=False):
    cplx = np.iscomplexobj(lu.factors)
    dt = np.complex128 if cplx else np.float64
    x = np.array(b, dtype=dt, order="F", copy=True)
    if x.ndim == 1:
        x = x.reshape(-1, 1, order="F")
    n = lu.factors.shape[0]
    piv = np.ascontiguousarray(lu.pivots, dtype=np.int32)
    fn = lib().or_zgetrs if cplx else lib().or_dgetrs
    fn(int(trans), n, x.shape[1], _ptr(lu.factors), max(1, n), _ptr(piv), _ptr(x), max(1, n))
    return x


def trsv(kind, l, b):
    l = _f(l, np.float64)
    x = np.array(b, dtype=np.float64, order="F", copy=True)
    m = l.shape[0]
    lib().or_dtrsv(int(kind), m, x.shape[1], _ptr(l), max(1, m), _ptr(x), max(1, m))
    return x


def gecon(lu: LU, columns=2):
    cplx = np.iscomplexobj(lu.factors)
    n = lu.factors.shape[0]
    piv = np.ascontiguousarray(lu.pivots, dtype=np.int32)
    fn = lib().or_zgecon if cplx else lib().or_dgecon
    return fn(n, _ptr(lu.factors), max(1, n), _ptr(piv), lu.singular_column, lu.anorm, columns)


# ---- estimators --------------------------------------------------------------
def normest_power(a, power=1, columns=2, max_iterations=5, seed=20260915):
    cplx = np.iscomplexobj(a)
    a = _f(a, np.complex128 if cplx else np.float64)
    n = a.shape[0]
    fn = lib().or_znormest_power if cplx else lib().or_dnormest_power
    r = fn(n, _ptr(a), max(1, n), power, columns, max_iterations, seed)
    return r.value, r.iterations, r.products


def normest_op(n, apply, adjoint, complex_=False, columns=2, max_iterations=5, seed=20260915):
    """Estimate ||op||_1 for Python callables apply(X)->Y, adjoint(X)->Y on n x t panels."""
    dt = np.complex128 if complex_ else np.float64

    def wrap(fn):
        def cb(_ctx, t, x, y):
            X = np.ctypeslib.as_array(C.cast(x, C.POINTER(C.c_double)), shape=(n * t * (2 if complex_ else 1),))
            X = X.view(dt).reshape((n, t), order="F")
            Y = np.ctypeslib.as_array(C.cast(y, C.POINTER(C.c_double)), shape=(n * t * (2 if complex_ else 1),))
            Y = Y.view(dt).reshape((n, t), order="F")
            Y[...] = fn(X.copy())
        return (APPLY_Z if complex_ else APPLY_D)(cb)

    ap, ad = wrap(apply), wrap(adjoint)
    fn = lib().or_znormest if complex_ else lib().or_dnormest
    r = fn(n, ap, ad, None, columns, max_iterations, seed)
    return r.value, r.iterations, r.products


class Random:
    """System.Random(seed) restated: Sample(), Next(max), NextDouble()."""

    def __init__(self, seed: int):
        self._buf = C.create_string_buffer(58 * 4)
        lib().or_random_init(self._buf, int(seed))

    def sample(self) -> float:
        return lib().or_random_sample(self._buf)

    next_double = sample

    def next(self, max_value: int) -> int:
        return lib().or_random_next(self._buf, int(max_value))


def sort_keys_items(keys, items):
    k = np.ascontiguousarray(keys, dtype=np.float64).copy()
    v = np.ascontiguousarray(items, dtype=np.int32).copy()
    lib().or_sort_keys_items(len(k), _ptr(k), _ptr(v))
    return k, v


# ---- exponentials ------------------------------------------------------------
def expm(a):
    cplx = np.iscomplexobj(a)
    dt = np.complex128 if cplx else np.float64
    a = _f(a, dt)
    n = a.shape[0]
    e = np.zeros((n, n), dtype=dt, order="F")
    d = ExpmDiag()
    fn = lib().or_zexpm if cplx else lib().or_dexpm
    rc = fn(n, _ptr(a), max(1, n), _ptr(e), max(1, n), C.byref(d))
    if rc != 0:
        raise ArithmeticError("Pade denominator is exactly singular")
    return e, (d.degree, d.squarings)


def expmv(a, b, t=1.0):
    cplx = np.iscomplexobj(a) or np.iscomplexobj(b)
    dt = np.complex128 if cplx else np.float64
    a = _f(a, dt); b = _f(b, dt)
    if b.ndim == 1:
        b = b.reshape(-1, 1, order="F")
    n, nb = b.shape
    f = np.zeros((n, nb), dtype=dt, order="F")
    d = ExpmvDiag()
    fn = lib().or_zexpmv if cplx else lib().or_dexpmv
    rc = fn(n, nb, _ptr(a), max(1, n), _ptr(b), max(1, n), float(t), _ptr(f), max(1, n), C.byref(d))
    if rc != 0:
        raise OverflowError("selected cost is not finite")
    return f, (d.degree, d.scaling, d.applications, d.early_exits)


def expmv_op(n, apply, b, t, bound, complex_=False):
    dt = np.complex128 if complex_ else np.float64
    b = _f(b, dt)
    nb = b.shape[1]

    def cb(_ctx, tt, x, y):
        w = 2 if complex_ else 1
        X = np.ctypeslib.as_array(C.cast(x, C.POINTER(C.c_double)), shape=(n * tt * w,)).view(dt).reshape((n, tt), order="F")
        Y = np.ctypeslib.as_array(C.cast(y, C.POINTER(C.c_double)), shape=(n * tt * w,)).view(dt).reshape((n, tt), order="F")
        Y[...] = apply(X.copy())

    f = np.zeros((n, nb), dtype=dt, order="F")
    d = ExpmvDiag()
    fnp = (APPLY_Z if complex_ else APPLY_D)(cb)
    fn = lib().or_zexpmv_op if complex_ else lib().or_dexpmv_op
    rc = fn(n, fnp, None, nb, _ptr(b), max(1, n), float(t), float(bound), _ptr(f), max(1, n), C.byref(d))
    if rc == -3:
        raise ValueError("norm bound must be finite and not negative")
    if rc != 0:
        raise OverflowError("selected cost is not finite")
    return f, (d.degree, d.scaling, d.applications, d.early_exits)


def expm_taylor_ld(a):
    cplx = np.iscomplexobj(a)
    dt = np.complex128 if cplx else np.float64
    a = _f(a, dt)
    n = a.shape[0]
    e = np.zeros((n, n), dtype=dt, order="F")
    fn = lib().or_zexpm_taylor_ld if cplx else lib().or_dexpm_taylor_ld
    fn(n, _ptr(a), max(1, n), _ptr(e), max(1, n))
    return e


def expmv_taylor_ld(a, b, t):
    """exp(tA) B with every operation in x87 extended precision."""
    a = _f(a, np.complex128)
    b = _f(np.asarray(b).reshape(a.shape[0], -1), np.complex128)
    n, nb = b.shape
    f = np.zeros((n, nb), dtype=np.complex128, order="F")
    lib().or_zexpmv_taylor_ld(n, _ptr(a), max(1, n), float(t), nb, _ptr(b), max(1, n), _ptr(f), max(1, n))
    return f


def cabs(z):
    return lib().or_cabs(_cplx(z))


def cdiv(x, y):
    r = lib().or_cdiv(_cplx(x), _cplx(y))
    return complex(r.re, r.im)
