"""Python mirror of Tensile's public surface for the dense hot path, over the
MI355X C ABI (libtensile_hip.so).

Names and argument meanings follow the reference's C# API
(src/Tensile/MatrixOperations.cs, LuDecomposition.cs, NormEstimate.cs,
MatrixExponential.cs, MatrixExponentialAction.cs, LinearOperators.cs);
failures raise the counterparts of its exceptions. Matrices are numpy arrays
(host memory, any order: they are passed column-major) or DeviceMatrix views
of device memory. Nothing here computes: every call goes to the library.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib as L


# ---- exceptions (the reference's exception types) ----------------------------
class TensileError(Exception):
    """Base of the library's errors."""


class ArgumentError(TensileError, ValueError):
    """ArgumentException: operands disagree."""


class ArgumentOutOfRangeError(TensileError, ValueError):
    """ArgumentOutOfRangeException."""


class InvalidOperationError(TensileError):
    """InvalidOperationException: singular factorization, singular Pade denominator."""


class AllocationLimitError(TensileError, MemoryError):
    """AllocationLimitException: the device working-memory ceiling."""


class ObjectDisposedError(TensileError):
    """ObjectDisposedException: the workspace was disposed."""


class DeviceError(TensileError, RuntimeError):
    """The HIP runtime failed."""


class OperatorError(TensileError):
    """A caller's linear operator raised; the original exception is chained."""


_ERRORS = {
    L.TNS_EINVAL: ArgumentError,
    L.TNS_ERANGE: ArgumentOutOfRangeError,
    L.TNS_ESTATE: InvalidOperationError,
    L.TNS_ENOMEM: AllocationLimitError,
    L.TNS_EDISPOSED: ObjectDisposedError,
    L.TNS_EDEVICE: DeviceError,
    L.TNS_ECALLBACK: OperatorError,
}


# ---- operands ---------------------------------------------------------------
@dataclass
class DeviceMatrix:
    """A column-major matrix in device memory: base address, shape, leading
    dimension and element type. `owner` keeps the allocation alive."""

    ptr: int
    rows: int
    cols: int
    ld: int
    complex_: bool = False
    owner: object = None

    @staticmethod
    def from_torch(t, rows: Optional[int] = None):
        """View a contiguous (cols, rows) torch tensor as a rows x cols matrix."""
        import torch  # local: the library itself does not need torch

        assert t.is_contiguous() and t.dim() == 2
        cplx = t.dtype == torch.complex128
        assert cplx or t.dtype == torch.float64
        cols, r = t.shape
        return DeviceMatrix(t.data_ptr(), rows if rows is not None else r, cols, r, cplx, t)


def _is_complex(*xs):
    for x in xs:
        if isinstance(x, DeviceMatrix):
            if x.complex_:
                return True
        elif x is not None and np.iscomplexobj(x):
            return True
    return False


class _Operand:
    """Address/shape of a host (numpy) or device operand."""

    def __init__(self, x, cplx: bool, vector_ok=True):
        if isinstance(x, DeviceMatrix):
            self.ptr, self.rows, self.cols, self.ld = x.ptr, x.rows, x.cols, x.ld
            self.array = None
        else:
            arr = np.asarray(x)
            if arr.ndim == 1 and vector_ok:
                arr = arr.reshape(-1, 1)
            if arr.ndim != 2:
                raise ArgumentError("operands must be matrices")
            arr = np.asfortranarray(arr, dtype=np.complex128 if cplx else np.float64)
            self.array = arr
            self.rows, self.cols = arr.shape
            self.ld = max(1, self.rows)
            self.ptr = arr.ctypes.data if arr.size else None


def _host_out(rows, cols, cplx):
    return np.zeros((rows, cols), dtype=np.complex128 if cplx else np.float64, order="F")


def _cplx(z) -> L.Complex:
    z = complex(z)
    return L.Complex(z.real, z.imag)


# ---- workspace ----------------------------------------------------------------
class Workspace:
    """The device context: stream, device-memory pool and limit
    (the device side of Tensile's Workspace, Workspace.cs:36-268)."""

    _shared = None

    def __init__(self, device: int = 0):
        self._lib = L.load()
        h = C.c_void_p()
        rc = self._lib.tns_context_create(int(device), C.byref(h))
        if rc != L.TNS_OK:
            raise _ERRORS.get(rc, DeviceError)(f"could not create a device context (status {rc})")
        self._h = h
        self._pending = None

    @classmethod
    def shared(cls) -> "Workspace":
        if cls._shared is None:
            cls._shared = Workspace()
        return cls._shared

    @property
    def handle(self):
        return self._h

    def check(self, rc: int):
        if rc == L.TNS_OK:
            return
        msg = self._lib.tns_last_error(self._h)
        msg = msg.decode() if msg else f"status {rc}"
        err = _ERRORS.get(rc, DeviceError)(msg)
        if rc == L.TNS_ECALLBACK and self._pending is not None:
            pending, self._pending = self._pending, None
            raise err from pending
        raise err

    def dispose(self):
        if self._h is not None:
            self._lib.tns_context_destroy(self._h)

    def set_limit(self, nbytes: int):
        self.check(self._lib.tns_context_set_limit(self._h, int(nbytes)))

    def set_retention(self, nbytes: int):
        """Cap on the free device memory kept between calls (the retention
        cap of ComplexScratch, ComplexScratch.cs:12-60)."""
        self.check(self._lib.tns_context_set_retention(self._h, int(nbytes)))

    def memory(self):
        """(held, limit, retention) of the context's device pool, in bytes."""
        held, limit, ret = C.c_size_t(), C.c_size_t(), C.c_size_t()
        self.check(self._lib.tns_context_memory(self._h, C.byref(held), C.byref(limit), C.byref(ret)))
        return held.value, limit.value, ret.value

    def synchronize(self):
        self.check(self._lib.tns_synchronize(self._h))

    def stream(self) -> int:
        s = C.c_void_p()
        self.check(self._lib.tns_context_stream(self._h, C.byref(s)))
        return s.value or 0

    # ---- GEMM (Workspace.Multiply, Workspace.cs:117-140) -------------------------
    def multiply(self, a, b, c=None, alpha=1.0, beta=0.0):
        """C := beta*C + alpha*A*B. With c=None a new result is returned."""
        cplx = _is_complex(a, b, c) or isinstance(alpha, complex) or isinstance(beta, complex)
        A, B = _Operand(a, cplx), _Operand(b, cplx)
        if A.cols != B.rows:
            raise ArgumentError(f"Inner dimensions disagree: A is {A.rows}x{A.cols}, B is {B.rows}x{B.cols}.")
        if c is None:
            c = _host_out(A.rows, B.cols, cplx)
            beta = 0.0
        if isinstance(c, DeviceMatrix):
            Cm = _Operand(c, cplx)
        else:
            if not (isinstance(c, np.ndarray) and c.flags.f_contiguous and c.dtype == (np.complex128 if cplx else np.float64)):
                raise ArgumentError("the destination must be a column-major array of the operands' type")
            Cm = _Operand(c, cplx)
            Cm.array = c
            Cm.ptr = c.ctypes.data if c.size else None
        if Cm.rows != A.rows or Cm.cols != B.cols:
            raise ArgumentError(f"Destination is {Cm.rows}x{Cm.cols}, expected {A.rows}x{B.cols}.")
        if cplx:
            rc = self._lib.tns_zgemm(self._h, A.rows, B.cols, A.cols, _cplx(alpha), A.ptr, A.ld, B.ptr, B.ld,
                                     _cplx(beta), Cm.ptr, Cm.ld)
        else:
            rc = self._lib.tns_dgemm(self._h, A.rows, B.cols, A.cols, float(alpha), A.ptr, A.ld, B.ptr, B.ld,
                                     float(beta), Cm.ptr, Cm.ld)
        self.check(rc)
        return c

    # ---- LU (Workspace.FactorLu, Workspace.cs:208-233) -------------------------
    def factor_lu(self, a, block_size: int = 0, timings: "LuPhaseTimings" = None) -> "LuDecomposition":
        """Factor a copy of A (MatrixOperations.FactorLu clones, MatrixOperations.cs:76-84).
        With `timings`, the factorization's phase split is added to it
        (Workspace.FactorLu's LuPhaseTimings argument, Workspace.cs:208-233)."""
        cplx = _is_complex(a)
        f = np.array(a, dtype=np.complex128 if cplx else np.float64, order="F", copy=True)
        if f.ndim != 2:
            raise ArgumentError("LU requires a matrix")
        m, n = f.shape
        anorm = self.one_norm(f) if f.size else 0.0
        piv = np.zeros(max(1, min(m, n)), dtype=np.int32)
        info = L.LuInfo()
        args = (self._h, m, n, f.ctypes.data if f.size else None, max(1, m), int(block_size), piv.ctypes.data,
                C.byref(info))
        if timings is None:
            fn = self._lib.tns_zgetrf if cplx else self._lib.tns_dgetrf
            self.check(fn(*args))
        else:
            fn = self._lib.tns_zgetrf_timed if cplx else self._lib.tns_dgetrf_timed
            self.check(fn(*args, C.byref(timings._raw)))
        return LuDecomposition(self, f, piv[: min(m, n)].copy(), info.singular_column, info.smallest_pivot,
                               info.largest_pivot, anorm)

    def factor_lu_device(self, a: "DeviceMatrix", pivots: "DeviceMatrix" = None, block_size: int = 0,
                         timings: "LuPhaseTimings" = None):
        """Factor a device matrix in place (KernelEntry.FactorLu on device
        memory). Returns (host pivots, singular column)."""
        cplx = a.complex_
        m, n = a.rows, a.cols
        piv = np.zeros(max(1, min(m, n)), dtype=np.int32)
        info = L.LuInfo()
        args = (self._h, m, n, a.ptr, a.ld, int(block_size), piv.ctypes.data, C.byref(info))
        if timings is None:
            fn = self._lib.tns_zgetrf if cplx else self._lib.tns_dgetrf
            self.check(fn(*args))
        else:
            fn = self._lib.tns_zgetrf_timed if cplx else self._lib.tns_dgetrf_timed
            self.check(fn(*args, C.byref(timings._raw)))
        return piv[: min(m, n)].copy(), info.singular_column

    # ---- norms (MatrixOperations.cs:142-174) ----------------------------------------
    def _norm(self, a, kind):
        cplx = _is_complex(a)
        A = _Operand(a, cplx)
        out = C.c_double()
        fn = self._lib.tns_znorm if cplx else self._lib.tns_dnorm
        self.check(fn(self._h, kind.encode(), A.rows, A.cols, A.ptr, A.ld, C.byref(out)))
        return out.value

    def one_norm(self, a) -> float:
        return self._norm(a, "1")

    def infinity_norm(self, a) -> float:
        return self._norm(a, "I")

    def frobenius_norm(self, a) -> float:
        return self._norm(a, "F")

    # ---- triangular solves and panels (KernelEntry.cs:98-201) -------------------------
    def solve_triangular(self, kind: int, a, b):
        A, B = _Operand(a, False), _Operand(b, False)
        x = np.array(B.array, order="F", copy=True)
        self.check(self._lib.tns_dtrsm(self._h, int(kind), A.rows, x.shape[1], A.ptr, A.ld,
                                       x.ctypes.data if x.size else None, max(1, A.rows)))
        return x

    def multiply_panel(self, a, x, transposed=False):
        cplx = _is_complex(a, x)
        A, X = _Operand(a, cplx), _Operand(x, cplx)
        y = _host_out(A.rows, X.cols, cplx)
        fn = self._lib.tns_zpanel if cplx else self._lib.tns_dpanel
        self.check(fn(self._h, int(bool(transposed)), A.rows, X.cols, A.ptr, A.ld, X.ptr, X.ld,
                      y.ctypes.data if y.size else None, max(1, A.rows)))
        return y

    # ---- estimators ---------------------------------------------------------------------
    def estimate_one_norm(self, a, power=1, columns=2, max_iterations=5, seed=20260915) -> "NormEstimateResult":
        cplx = _is_complex(a)
        A = _Operand(a, cplx)
        if A.rows != A.cols:
            raise ArgumentError(f"Norm estimation requires a square matrix, got {A.rows}x{A.cols}.")
        out = L.NormEst()
        fn = self._lib.tns_znormest_power if cplx else self._lib.tns_dnormest_power
        self.check(fn(self._h, A.rows, A.ptr, A.ld, int(power), int(columns), int(max_iterations), int(seed),
                      C.byref(out)))
        return NormEstimateResult(out.value, out.iterations, out.products)

    # ---- exponentials ---------------------------------------------------------------------
    def batched_systems(self, matrices, rhs, tune: int = 1) -> "BatchedSystems":
        """A BatchedSystems plan over device matrices (see that class)."""
        return BatchedSystems(self, matrices, rhs, tune)

    def solve_mixed(self, a, b, precision: int = 32, tol: float = 1e-15, max_iterations: int = 30, out=None):
        """X with A X = B by a low-precision LU (FP32, or BF16 products with
        equilibration when precision=16) and FP64 iterative refinement
        (FGMRES preconditioned by the low-precision solve, at most
        max_iterations solves) to a backward error <= tol, falling back to the
        FP64 LU when refinement stalls or would cost more than it. Returns
        (X, MixedInfo). Real matrices only."""
        if _is_complex(a, b):
            raise ArgumentError("The mixed-precision solve takes real matrices.")
        A = _Operand(a, False)
        B = _Operand(b, False)
        if A.rows != A.cols or B.rows != A.rows:
            raise ArgumentError(f"Shapes disagree: A is {A.rows}x{A.cols}, B is {B.rows}x{B.cols}.")
        n, t = A.rows, B.cols
        X = out if out is not None else _host_out(n, t, False)
        Xo = _Operand(X, False) if isinstance(X, DeviceMatrix) else None
        xptr, xld = (Xo.ptr, Xo.ld) if Xo else (X.ctypes.data if X.size else None, max(1, n))
        info = L.MixedInfo()
        self.check(self._lib.tns_dgesv_mixed(self._h, n, t, A.ptr, A.ld, B.ptr, B.ld, xptr, xld, int(precision),
                                             float(tol), int(max_iterations), C.byref(info)))
        return X, MixedInfo(info.precision, info.iterations, bool(info.fell_back), info.backward_error,
                            info.factor_ms, info.refine_ms, info.fallback_ms, info.first_error)

    def expm(self, a, out=None, phases: "ExpmPhaseTimings" = None):
        """exp(A) and its ExpmDiagnostics (MatrixExponential.cs:143-270).
        phases, if given, receives this call's device-time split (powers,
        degree/scaling selection, Pade products, LU, solve, squarings)."""
        cplx = _is_complex(a)
        A = _Operand(a, cplx)
        if A.rows != A.cols:
            raise ArgumentError(f"The matrix exponential requires a square matrix, got {A.rows}x{A.cols}.")
        n = A.rows
        E = out if out is not None else _host_out(n, n, cplx)
        Eo = _Operand(E, cplx) if isinstance(E, DeviceMatrix) else None
        eptr, eld = (Eo.ptr, Eo.ld) if Eo else (E.ctypes.data if E.size else None, max(1, n))
        d = L.ExpmDiag()
        if phases is None:
            fn = self._lib.tns_zexpm if cplx else self._lib.tns_dexpm
            self.check(fn(self._h, n, A.ptr, A.ld, eptr, eld, C.byref(d)))
        else:
            ph = L.ExpmPhases()
            fn = self._lib.tns_zexpm_timed if cplx else self._lib.tns_dexpm_timed
            self.check(fn(self._h, n, A.ptr, A.ld, eptr, eld, C.byref(d), C.byref(ph)))
            phases.add(ph)
        return E, ExpmDiagnostics(d.degree, d.squarings)

    def expmv(self, a, b, t=1.0, out=None):
        """exp(tA) B and its ExpmvDiagnostics (MatrixExponentialAction.cs:102-178)."""
        cplx = _is_complex(a, b)
        A, B = _Operand(a, cplx), _Operand(b, cplx)
        if A.rows != A.cols:
            raise ArgumentError(f"The exponential action requires a square matrix, got {A.rows}x{A.cols}.")
        if B.rows != A.rows:
            raise ArgumentError(f"B has {B.rows} rows, expected the operator's order {A.rows}.")
        F = out if out is not None else _host_out(B.rows, B.cols, cplx)
        Fo = _Operand(F, cplx) if isinstance(F, DeviceMatrix) else None
        fptr, fld = (Fo.ptr, Fo.ld) if Fo else (F.ctypes.data if F.size else None, max(1, B.rows))
        d = L.ExpmvDiag()
        fn = self._lib.tns_zexpmv if cplx else self._lib.tns_dexpmv
        self.check(fn(self._h, A.rows, B.cols, A.ptr, A.ld, B.ptr, B.ld, float(t), fptr, fld, C.byref(d)))
        return F, ExpmvDiagnostics(d.degree, d.scaling, d.applications, d.early_exits)

    def expmv_backward(self, a, b, t, grad_f, out_a=None, out_b=None):
        """The gradients (dL/dA, dL/dB) through F = exp(tA) B for an
        upstream gradient grad_f = dL/dF, by the adjoint of the Taylor
        algorithm with recomputed stages (see ExpmvFunction)."""
        cplx = _is_complex(a, b, grad_f)
        A, B, G = _Operand(a, cplx), _Operand(b, cplx), _Operand(grad_f, cplx)
        n, nb = A.rows, B.cols
        if A.rows != A.cols or B.rows != n or G.rows != n or G.cols != nb:
            raise ArgumentError("Shapes disagree: A n x n, B and dL/dF n x nb.")
        GA = out_a if out_a is not None else _host_out(n, n, cplx)
        GB = out_b if out_b is not None else _host_out(n, nb, cplx)
        def ptr(x, rows):
            if isinstance(x, DeviceMatrix):
                o = _Operand(x, cplx)
                return o.ptr, o.ld
            return (x.ctypes.data if x.size else None), max(1, rows)
        pa, la = ptr(GA, n)
        pb, lb = ptr(GB, n)
        d = L.ExpmvDiag()
        fn = self._lib.tns_zexpmv_grad if cplx else self._lib.tns_dexpmv_grad
        self.check(fn(self._h, n, nb, A.ptr, A.ld, B.ptr, B.ld, float(t), G.ptr, G.ld, pa, la, pb, lb, C.byref(d)))
        return GA, GB

    def expmv_operator(self, op: "ILinearOperator", b, t: float, one_norm_bound: float, complex_=None):
        """The matrix-free action over a Python ILinearOperator
        (MatrixExponentialAction.cs:203-230)."""
        cplx = _is_complex(b) if complex_ is None else complex_
        dt = np.complex128 if cplx else np.float64
        B = _Operand(b, cplx)
        n = op.order
        if B.rows != n:
            raise ArgumentError(f"B has {B.rows} rows, expected the operator's order {n}.")
        F = _host_out(B.rows, B.cols, cplx)
        width = 2 if cplx else 1

        def cb(_user, nn, tt, x, ldx, y, ldy):
            try:
                X = np.ctypeslib.as_array(C.cast(x, C.POINTER(C.c_double)), shape=(ldx * tt * width,))
                X = X.view(dt).reshape((ldx, tt), order="F")[:nn]
                Y = np.ctypeslib.as_array(C.cast(y, C.POINTER(C.c_double)), shape=(ldy * tt * width,))
                Y = Y.view(dt).reshape((ldy, tt), order="F")
                Y[:nn] = op.apply(X.copy())
                return 0
            except BaseException as e:  # surfaced by check() as OperatorError
                self._pending = e
                return 1

        fnp = L.ZAPPLY(cb) if cplx else L.DAPPLY(cb)
        d = L.ExpmvDiag()
        fn = self._lib.tns_zexpmv_op if cplx else self._lib.tns_dexpmv_op
        self.check(fn(self._h, n, fnp, None, 0, B.cols, B.ptr, B.ld, float(t), float(one_norm_bound),
                      F.ctypes.data if F.size else None, max(1, n), C.byref(d)))
        return F, ExpmvDiagnostics(d.degree, d.scaling, d.applications, d.early_exits)

    def expmv_device_operator(self, op: "DeviceLinearOperator", b, t: float, one_norm_bound: float, out=None):
        """The matrix-free action over a device-resident operator
        (MatrixExponentialAction.cs:203-230 with ILinearOperator<T>,
        LinearOperators.cs:37-75, applied to panels in device memory): every
        application hands op.apply_device two DeviceMatrix panels and the
        context's stream; the operator's work must be complete, or ordered on
        that stream, when it returns. B and the result may be host or device
        matrices."""
        cplx = bool(op.complex_)
        B = _Operand(b, cplx)
        n = op.order
        if B.rows != n:
            raise ArgumentError(f"B has {B.rows} rows, expected the operator's order {n}.")
        F = out if out is not None else _host_out(B.rows, B.cols, cplx)
        Fo = _Operand(F, cplx) if isinstance(F, DeviceMatrix) else None
        fptr, fld = (Fo.ptr, Fo.ld) if Fo else (F.ctypes.data if F.size else None, max(1, n))
        stream = self.stream()

        def cb(_user, nn, tt, x, ldx, y, ldy):
            try:
                op.apply_device(DeviceMatrix(int(x), nn, tt, ldx, cplx), DeviceMatrix(int(y), nn, tt, ldy, cplx),
                                stream)
                return 0
            except BaseException as e:  # surfaced by check() as OperatorError
                self._pending = e
                return 1

        fnp = L.ZAPPLY(cb) if cplx else L.DAPPLY(cb)
        d = L.ExpmvDiag()
        fn = self._lib.tns_zexpmv_op if cplx else self._lib.tns_dexpmv_op
        self.check(fn(self._h, n, fnp, None, 1, B.cols, B.ptr, B.ld, float(t), float(one_norm_bound), fptr, fld,
                      C.byref(d)))
        return F, ExpmvDiagnostics(d.degree, d.scaling, d.applications, d.early_exits)


# ---- results --------------------------------------------------------------------
@dataclass(frozen=True)
class NormEstimateResult:
    value: float
    iterations: int
    products: int


@dataclass(frozen=True)
class MixedInfo:
    precision: int
    iterations: int
    fell_back: bool
    backward_error: float
    factor_ms: float
    refine_ms: float
    fallback_ms: float
    first_error: float


@dataclass(frozen=True)
class ExpmDiagnostics:
    degree: int
    squarings: int


@dataclass(frozen=True)
class ExpmvDiagnostics:
    degree: int
    scaling: int
    applications: int
    early_exits: int


class ExpmvFunction:
    """exp(tA) B as a differentiable function in the style of an autograd
    Function: forward(a, b, t) returns F and keeps its inputs; backward(grad_f)
    returns (dL/dA, dL/dB) for grad_f = dL/dF. For complex data the gradients
    g satisfy dL = Re tr(g^H dA) (the real-loss convention)."""

    def __init__(self, workspace: "Workspace" = None):
        self._ws = workspace or Workspace.shared()
        self._saved = None

    def forward(self, a, b, t: float = 1.0):
        f, diag = self._ws.expmv(a, b, t)
        self._saved = (a, b, float(t))
        self.diagnostics = diag
        return f

    def backward(self, grad_f):
        if self._saved is None:
            raise InvalidOperationError("backward() before forward()")
        a, b, t = self._saved
        return self._ws.expmv_backward(a, b, t, grad_f)


class BatchedSystems:
    """Many complex systems Z_i X_i = B_i of uneven order, factored and
    solved in one launch of a captured hipGraph of grouped kernels, each
    block step of every system's LU in one launch (the frequency sweep). The
    matrices must be device memory (DeviceMatrix); solve() overwrites each
    Z_i with its LU factors and each B_i with X_i. tune: 0 defaults, 1 the
    workspace's remembered choice for this batch shape (tuning on scratch
    copies if there is none), 2 tune again."""

    def __init__(self, ws: "Workspace", matrices, rhs, tune: int = 1):
        if len(matrices) != len(rhs):
            raise ArgumentError("one right-hand side block per matrix")
        self._ws = ws
        self._h = None
        self.count = len(matrices)
        for m, r in zip(matrices, rhs):
            if not isinstance(m, DeviceMatrix) or not isinstance(r, DeviceMatrix):
                raise ArgumentError("batched systems must be DeviceMatrix operands")
        nrhs = rhs[0].cols if rhs else 0
        n = np.array([m.rows for m in matrices], dtype=np.int32)
        lda = np.array([m.ld for m in matrices], dtype=np.int32)
        ldb = np.array([r.ld for r in rhs], dtype=np.int32)
        for m, r in zip(matrices, rhs):
            if m.rows != m.cols or r.rows != m.rows or r.cols != nrhs:
                raise ArgumentError("each matrix square, each right-hand side n_i x nrhs")
        a = (C.c_void_p * max(1, self.count))(*[m.ptr for m in matrices])
        b = (C.c_void_p * max(1, self.count))(*[r.ptr for r in rhs])
        self._n = n
        self._keep = (n, lda, ldb, a, b, matrices, rhs)
        h = C.c_void_p()
        ws.check(ws._lib.tns_zbatch_create(ws.handle, self.count, n.ctypes.data, C.cast(a, C.c_void_p),
                                           lda.ctypes.data, C.cast(b, C.c_void_p), ldb.ctypes.data, nrhs,
                                           int(tune), C.byref(h)))
        self._h = h

    def solve(self) -> np.ndarray:
        """One graph launch; returns each system's first zero pivot or -1."""
        out = np.zeros(self.count, dtype=np.int32)
        self._ws.check(self._ws._lib.tns_zbatch_solve(self._h, out.ctypes.data if self.count else None))
        return out

    def reciprocal_condition(self, columns: int = 2) -> np.ndarray:
        out = np.zeros(self.count, dtype=np.float64)
        self._ws.check(self._ws._lib.tns_zbatch_rcond(self._h, int(columns), out.ctypes.data if self.count else
                                                       C.cast(C.pointer(C.c_double()), C.c_void_p)))
        return out

    def pivots(self, index: int) -> np.ndarray:
        out = np.zeros(max(1, int(self._n[index])), dtype=np.int32)
        self._ws.check(self._ws._lib.tns_zbatch_pivots(self._h, int(index), out.ctypes.data))
        return out[: int(self._n[index])]

    @property
    def config(self):
        """(LU panel width, grouped GEMM tile, the tuner's best graph ms)."""
        s, g, t = C.c_int(), C.c_int(), C.c_double()
        self._ws.check(self._ws._lib.tns_zbatch_info(self._h, C.byref(s), C.byref(g), C.byref(t)))
        return s.value, g.value, t.value

    def close(self):
        if self._h:
            self._ws._lib.tns_zbatch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LuDecomposition:
    """LuDecomposition<T> (LuDecomposition.cs:37-237): the factors, pivots,
    diagnostics and ||A||_1 captured before factoring."""

    def __init__(self, ws: Workspace, factors, pivots, singular_column, smallest, largest, one_norm_of_a):
        self._ws = ws
        self.factors = factors
        self.pivots = pivots
        self.singular_column = int(singular_column)
        self.smallest_pivot = float(smallest)
        self.largest_pivot = float(largest)
        self.one_norm_of_a = float(one_norm_of_a)

    @property
    def rows(self):
        return self.factors.shape[0]

    @property
    def columns(self):
        return self.factors.shape[1]

    @property
    def is_square(self):
        return self.rows == self.columns

    @property
    def is_singular(self):
        return self.singular_column >= 0

    @property
    def pivot_ratio(self):
        return 0.0 if self.largest_pivot == 0.0 else self.smallest_pivot / self.largest_pivot

    def _solve(self, b, trans):
        if not self.is_square:
            raise InvalidOperationError("Solve requires a square factorization.")
        cplx = np.iscomplexobj(self.factors)
        x = np.array(b, dtype=np.complex128 if cplx else np.float64, order="F", copy=True)
        vec = x.ndim == 1
        if vec:
            x = x.reshape(-1, 1, order="F")
        if x.shape[0] != self.rows:
            raise ArgumentError(f"Right-hand side has {x.shape[0]} rows, expected {self.rows}.")
        n = self.rows
        fn = self._ws._lib.tns_zgetrs if cplx else self._ws._lib.tns_dgetrs
        self._ws.check(fn(self._ws.handle, int(trans), n, x.shape[1], self.factors.ctypes.data if n else None,
                          max(1, n), self.pivots.ctypes.data if n else None, self.singular_column,
                          x.ctypes.data if x.size else None, max(1, n)))
        return x[:, 0] if vec else x

    def solve(self, b):
        """X with A X = B (LuDecomposition.Solve, LuDecomposition.cs:106-121)."""
        return self._solve(b, 0)

    def solve_adjoint(self, b):
        """X with A^H X = B (A^T for real)."""
        return self._solve(b, 1)

    def reciprocal_condition(self, columns: int = 2) -> float:
        """1/(||A||_1 ||A^-1||_1), estimated (MatrixOperations.cs:106-114)."""
        if not self.is_square:
            raise InvalidOperationError("Condition estimation requires a square factorization.")
        cplx = np.iscomplexobj(self.factors)
        n = self.rows
        out = C.c_double()
        fn = self._ws._lib.tns_zgecon if cplx else self._ws._lib.tns_dgecon
        self._ws.check(fn(self._ws.handle, n, self.factors.ctypes.data if n else None, max(1, n),
                          self.pivots.ctypes.data if n else None, self.singular_column, self.one_norm_of_a,
                          int(columns), C.byref(out)))
        return out.value

    def determinant(self):
        """Product of U's diagonal with the permutation's sign (LuDecomposition.cs:139-157).
        A host-side O(n) reduction over the factors."""
        if not self.is_square:
            raise InvalidOperationError("Determinant requires a square factorization.")
        prod = (1 + 0j) if np.iscomplexobj(self.factors) else 1.0
        for i in range(self.rows):
            prod *= self.factors[i, i]
        for k, p in enumerate(self.pivots):
            if p != k:
                prod = -prod
        return prod


class ILinearOperator:
    """ILinearOperator<T> (LinearOperators.cs:37-57): a square operator that
    can be applied to an n x t panel. Subclasses set `order` and `apply`."""

    order: int = 0

    def apply(self, x):  # pragma: no cover - interface
        raise NotImplementedError


class DeviceLinearOperator:
    """ILinearOperator<T> whose panels live in device memory
    (LinearOperators.cs:37-75): apply_device(x, y, stream) writes Op*x into
    y, both DeviceMatrix panels of `order` rows; its work must be complete,
    or ordered on `stream` (the calling context's), when it returns.
    Subclasses set `order` and `complex_`."""

    order: int = 0
    complex_: bool = False

    def apply_device(self, x: DeviceMatrix, y: DeviceMatrix, stream: int):  # pragma: no cover - interface
        raise NotImplementedError


class DenseDeviceOperator(DeviceLinearOperator):
    """A device matrix as a device-resident operator: y := A x by this
    library's product on a second workspace (the calling context is busy
    with the action), synchronized before returning."""

    def __init__(self, a: DeviceMatrix, workspace: "Workspace"):
        if a.rows != a.cols:
            raise ArgumentError("the operator must be square")
        self.a = a
        self.order = a.rows
        self.complex_ = a.complex_
        self._ws = workspace

    def apply_device(self, x, y, stream):
        self._ws.multiply(self.a, x, y)
        self._ws.synchronize()


class LuPhaseTimings:
    """LuPhaseTimings (Lu.cs:92-164): device milliseconds charged to the
    panel factorizations, the row interchanges, the U12 triangular solves and
    the trailing GEMMs, the whole factorization's total, and the block steps;
    a factorization passed this object adds its split (pooling, as Add)."""

    def __init__(self):
        self._raw = L.LuPhases()

    panel = property(lambda self: self._raw.panel_ms)
    swaps = property(lambda self: self._raw.swaps_ms)
    triangular = property(lambda self: self._raw.trsm_ms)
    gemm = property(lambda self: self._raw.gemm_ms)
    total = property(lambda self: self._raw.total_ms)
    block_steps = property(lambda self: self._raw.block_steps)

    @property
    def unattributed(self) -> float:
        return max(0.0, self.total - (self.panel + self.swaps + self.triangular + self.gemm))

    def share(self, ms: float) -> float:
        """A phase's share of the total, as a percentage."""
        return 0.0 if self.total == 0 else 100.0 * ms / self.total

    def add(self, other: "LuPhaseTimings"):
        for f in ("panel_ms", "swaps_ms", "trsm_ms", "gemm_ms", "total_ms", "block_steps"):
            setattr(self._raw, f, getattr(self._raw, f) + getattr(other._raw, f))


class ExpmPhaseTimings:
    """Device milliseconds of matrix exponentials by phase: the powers A^2,
    A^4, A^6 and their norms; the degree and scaling selection (norm
    estimates, ell corrections); the Pade products and combinations; the
    denominator's LU; its solve; the squarings; and the total. An Expm given
    this object adds its split (pooling over calls)."""

    FIELDS = ("powers_ms", "select_ms", "pade_ms", "lu_ms", "solve_ms", "squaring_ms", "total_ms")

    def __init__(self):
        self.calls = 0
        for f in self.FIELDS:
            setattr(self, f, 0.0)

    def add(self, raw):
        self.calls += 1
        for f in self.FIELDS:
            setattr(self, f, getattr(self, f) + getattr(raw, f))

    def per_call(self) -> dict:
        c = max(1, self.calls)
        return {f[:-3]: round(getattr(self, f) / c, 3) for f in self.FIELDS}


# ---- fluent functions over the shared workspace (MatrixOperations.cs) ---------------------
def _ws(workspace):
    return workspace if workspace is not None else Workspace.shared()


def multiply(a, b, workspace=None):
    return _ws(workspace).multiply(a, b)


def multiply_into(a, b, destination, alpha=1.0, beta=0.0, workspace=None):
    return _ws(workspace).multiply(a, b, destination, alpha, beta)


def factor_lu(a, block_size=0, workspace=None, timings=None) -> LuDecomposition:
    return _ws(workspace).factor_lu(a, block_size, timings)


def solve(a, b, workspace=None):
    a = np.asarray(a)
    if a.shape[0] != a.shape[1]:
        raise ArgumentError(f"Solve requires a square matrix, got {a.shape[0]}x{a.shape[1]}.")
    return factor_lu(a, workspace=workspace).solve(b)


def one_norm(a, workspace=None):
    return _ws(workspace).one_norm(a)


def infinity_norm(a, workspace=None):
    return _ws(workspace).infinity_norm(a)


def frobenius_norm(a, workspace=None):
    return _ws(workspace).frobenius_norm(a)


def estimate_one_norm(a, power=1, columns=2, workspace=None) -> float:
    return _ws(workspace).estimate_one_norm(a, power, columns).value


def expm(a, workspace=None):
    return _ws(workspace).expm(a)[0]


def expmv(a, b, t=1.0, workspace=None):
    return _ws(workspace).expmv(a, b, t)[0]
