#!/usr/bin/env python3
"""Benchmark of the MI355X dense-path backend.

Workload (the step): one complex128 matrix exponential of order 1000,
A = -D + iW with ||A||_1 = 50 (degree-13 Pade, s squarings, 4M products on
FP64 matrix cores), inputs resident in device memory. `value` is its
algorithmic rate in GFLOP/s: ((6 + s) * 8n^3 + (8/3)n^3 + 8n^3) per step.

Also reported on the same line: the step's device-time split by phase
(HIP events between the phases of one timed Expm), the dominant kernel's
roofline (the 4M complex GEMM at n = 1000, HIP events around each launch on
the library's stream), the exponential action exp(tA)B of the same
operator (B 1000 x 8, t = 0.5), the FP64 LU + solve + rcond at n = 1024
(host operands and device-resident), the FP64 GEMM at n = 8192, the
mixed-precision solves against the FP64 LU (and against an FP32
partial-pivoting LU with FP64 refinement), the batched sweep, the Expmv
gradients, and beside each a CPU baseline: the oracle (a C restatement of
Tensile's algorithms, OpenMP GEMM) on a bounded sample of the same workload
on the host.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
  python bench.py --roofline-only   # the dominant kernel alone (20 launches)
                                    # and a 1 GiB device copy, for rocprofv3

`roofline.traffic` is read from profiles/zgemm_n1000_traffic.json when that
file exists: HBM bytes per launch from rocprofv3 --pmc FETCH_SIZE and
WRITE_SIZE runs of --roofline-only, scaled by what the same counters read
for the 1 GiB copy.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def to_device(x):
    import torch

    from csharp_linearalgebra_amd import DeviceMatrix

    t = torch.from_numpy(np.ascontiguousarray(np.asfortranarray(x).T)).cuda()
    return DeviceMatrix.from_torch(t)


def empty_device(rows, cols, complex_):
    import torch

    from csharp_linearalgebra_amd import DeviceMatrix

    t = torch.zeros((cols, rows), dtype=torch.complex128 if complex_ else torch.float64, device="cuda")
    return DeviceMatrix.from_torch(t)


def event_time(ws, fn, reps):
    """Average device time of fn() over reps launches, with HIP events
    recorded on the library's own stream."""
    import torch

    stream = torch.cuda.ExternalStream(ws.stream())
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    fn()
    ws.synchronize()
    start.record(stream)
    for _ in range(reps):
        fn()
    stop.record(stream)
    stop.synchronize()
    ws.synchronize()
    return start.elapsed_time(stop) / reps / 1e3


def launch_time(ws, fn, reps):
    """Average device time of one fn() launch: HIP events recorded on the
    library's stream right before and after each launch (what a kernel
    trace reports as the kernel's duration, plus the event overhead)."""
    import torch

    stream = torch.cuda.ExternalStream(ws.stream())
    fn()
    ws.synchronize()
    pairs = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        fn()
        b.record(stream)
        pairs.append((a, b))
    ws.synchronize()
    return sum(a.elapsed_time(b) for a, b in pairs) / reps / 1e3


def roofline_only():
    """The dominant kernel alone: 20 launches of the 4M ZGEMM at n = 1000 on
    the benchmark's operator, then 20 device copies of 1 GiB (the counter
    calibration: 1 GiB read and 1 GiB written per launch)."""
    import torch

    import csharp_linearalgebra_amd as T
    from csharp_linearalgebra_amd import workloads as W

    ws = T.Workspace(0)
    n = 1000
    a, _ = W.expm_operator(n, 50.0, 8, seed=0)
    dA = to_device(a)
    dC = empty_device(n, n, True)
    torch.cuda.synchronize()
    s = launch_time(ws, lambda: ws.multiply(dA, dA, dC), 20)
    src = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    src.fill_(1.0)
    torch.cuda.synchronize()
    for _ in range(20):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"zgemm_n1000_us": round(s * 1e6, 2), "copy_bytes": 2 * src.numel() * 8}))
    ws.dispose()


def mixed_leg(ws, n, kappas, nrhs=32):
    """The mixed-precision workload: A = U diag(sigma) V^T (64 Householder
    reflectors each side, sigma log-spaced to condition kappa), B n x 32
    Gaussian, on the device. Times the FP64 LU + solve (getrf + getrs on
    device copies) and the mixed solve in FP32 and BF16 (factorisation +
    FP64 refinement to backward error 1e-15, or the FP64 fallback)."""
    import ctypes as C

    import torch

    from csharp_linearalgebra_amd import DeviceMatrix
    from csharp_linearalgebra_amd import _lib as L
    from csharp_linearalgebra_amd import workloads as W

    lib = ws._lib
    out = []
    for kappa in kappas:
        a, b = W.conditioned_device(n, kappa, nrhs, seed=0)
        at = a.T.contiguous()  # column-major storage
        bt = b.T.contiguous()
        dA, dB = DeviceMatrix.from_torch(at), DeviceMatrix.from_torch(bt)
        x64 = torch.empty_like(bt)
        f = torch.empty_like(at)
        piv = torch.zeros(n, dtype=torch.int32, device="cuda")
        info = L.LuInfo()

        def fp64():
            f.copy_(at)
            x64.copy_(bt)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ws.check(lib.tns_dgetrf(ws.handle, n, n, f.data_ptr(), n, 0, piv.data_ptr(), C.byref(info)))
            ws.check(lib.tns_dgetrs(ws.handle, 0, n, nrhs, f.data_ptr(), n, piv.data_ptr(), info.singular_column,
                                    x64.data_ptr(), n))
            ws.synchronize()
            return time.perf_counter() - t0

        fp64()
        t64 = min(fp64() for _ in range(2))
        r = bt.T - a @ x64.T
        den = a.abs().sum(1).max() * x64.T.abs().max(0).values + bt.T.abs().max(0).values
        berr64 = float((r.abs().max(0).values / den).max())
        row = {"kappa": kappa, "fp64_ms": round(t64 * 1e3, 2), "fp64_backward_error": berr64}
        if kappa <= 1e2:
            row["fp32_partial_pivoting_ir"] = fp32_pp_ir(ws, a, at, bt, n, nrhs, t64)
        for prec in (32, 16):
            dX = DeviceMatrix.from_torch(torch.empty_like(bt))
            ws.solve_mixed(dA, dB, precision=prec, out=dX)
            best, inf = None, None
            for _ in range(2):
                ws.synchronize()
                t0 = time.perf_counter()
                _, inf = ws.solve_mixed(dA, dB, precision=prec, out=dX)
                ws.synchronize()
                dt = time.perf_counter() - t0
                best = dt if best is None else min(best, dt)
            key = "fp32" if prec == 32 else "bf16"
            row[key] = {"ms": round(best * 1e3, 2), "speedup": round(t64 / best, 3),
                        "backward_error": inf.backward_error, "iterations": inf.iterations,
                        "fell_back": inf.fell_back, "first_error": inf.first_error,
                        "factor_ms": round(inf.factor_ms, 2), "refine_ms": round(inf.refine_ms, 2),
                        "fallback_ms": round(inf.fallback_ms, 2)}
        out.append(row)
        del a, b, at, bt, f, x64
        torch.cuda.empty_cache()
    return out


def fp32_pp_ir(ws, a, at, bt, n, nrhs, t64, tol=1e-15, max_iter=30):
    """The precision gain apart from the random butterfly transform: FP32 LU
    with partial pivoting on the FP32 matrix cores (tns_sgetrf), then
    classical iterative refinement - residual and update in FP64, each
    correction through the FP32 factors (tns_sgetrs) - until the normwise
    backward error of every column is at most tol."""
    import ctypes as C

    import torch

    lib = ws._lib
    piv = np.zeros(n, dtype=np.int32)
    sc = C.c_int()
    anorm = a.abs().sum(1).max()

    def run():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f32 = at.float()
        torch.cuda.synchronize()
        ws.check(lib.tns_sgetrf(ws.handle, 32, n, f32.data_ptr(), n, piv.ctypes.data, C.byref(sc)))
        tf = time.perf_counter()
        x = torch.zeros_like(bt)
        r = bt.clone()
        it, berr = 0, float("inf")
        while it < max_iter:
            c = r.float().contiguous()
            torch.cuda.synchronize()
            ws.check(lib.tns_sgetrs(ws.handle, 32, n, nrhs, f32.data_ptr(), n, piv.ctypes.data, c.data_ptr(), n))
            x += c.double()
            it += 1
            r = bt - (a @ x.T).T
            den = anorm * x.abs().max(1).values + bt.abs().max(1).values
            berr = float((r.abs().max(1).values / den).max())
            if not berr > tol:
                break
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        return t1 - t0, tf - t0, it, berr

    run()
    dt, tf, it, berr = min(run() for _ in range(2))
    return {"ms": round(dt * 1e3, 2), "speedup": round(t64 / dt, 3), "factor_ms": round(tf * 1e3, 2),
            "iterations": it, "backward_error": berr, "converged": berr <= tol}


def sweep_leg(ws, count=512, nmin=64, nmax=1024, nrhs=4, cpu_sample=False):
    """The frequency sweep: `count` complex systems Z(w) = R + i w L of order
    n ~ U{nmin..nmax} with 4 right-hand sides, factored and solved in one
    graph launch (autotuned), then rcond of every system."""
    import torch

    from csharp_linearalgebra_amd import DeviceMatrix
    from csharp_linearalgebra_amd import workloads as W

    mats, rhs = W.frequency_sweep_device(count, nmin, nmax, nrhs, seed=0)
    work_a = [m.clone() for m in mats]
    work_b = [b.clone() for b in rhs]
    dm = [DeviceMatrix.from_torch(t) for t in work_a]
    db = [DeviceMatrix.from_torch(t) for t in work_b]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan = ws.batched_systems(dm, db, tune=1)
    tune_s = time.perf_counter() - t0
    best = None
    for _ in range(3):
        for w, m in zip(work_a, mats):
            w.copy_(m)
        for w, b in zip(work_b, rhs):
            w.copy_(b)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sing = plan.solve()
        ws.synchronize()
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    # rcond timed like the solve (best of three calls); the first call, which
    # also pays for the estimators' first use of fresh host memory, reported
    # beside it.
    rc_first = rc_s = None
    for _ in range(3):
        t0 = time.perf_counter()
        rc = plan.reciprocal_condition()
        dt = time.perf_counter() - t0
        rc_first = dt if rc_first is None else rc_first
        rc_s = dt if rc_s is None else min(rc_s, dt)
    block, cfg, tuned_ms = plan.config
    orders = [m.shape[0] for m in mats]
    # Residual check of every system (host): max backward error.
    worst = 0.0
    for m, b, x in zip(mats, rhs, work_b):
        a = m.T
        xx = x.T
        r = b.T - a @ xx
        den = a.abs().sum(1).max() * xx.abs().max() + b.abs().max()
        worst = max(worst, float(r.abs().max() / den))
    plan.close()
    cpu = None
    if cpu_sample:
        import oracle as O

        # The first 8 systems of the same sweep on the host: LU, solve, rcond.
        sample = list(zip(mats, rhs))[:8]
        hosts = [(m.T.cpu().numpy(), b.T.cpu().numpy()) for m, b in sample]
        t0 = time.perf_counter()
        for a_h, b_h in hosts:
            f = O.getrf(a_h)
            O.getrs(f, b_h)
            O.gecon(f)
        cs = time.perf_counter() - t0
        cpu = {"value": round(len(hosts) / cs, 2), "unit": "solves+rcond/s", "cores": O.threads(), "kind": "port",
               "sample": "the sweep's first %d systems (orders %s): LU, 4-RHS solve and rcond each, %.2f s"
                         % (len(hosts), [int(m.shape[0]) for m, _ in sample], cs)}
    return {"workload": "%d complex128 systems, n ~ U{%d..%d}, %d RHS, one graph launch" % (count, nmin, nmax, nrhs),
            "solve_ms": round(best * 1e3, 2), "solves_per_s": round(count / best, 1),
            "gflops": round(W.sweep_flops(orders, nrhs) / best / 1e9, 1),
            "rcond_ms": round(rc_s * 1e3, 2), "rcond_first_call_ms": round(rc_first * 1e3, 2),
            "solves_plus_rcond_per_s": round(count / (best + rc_s), 1),
            "tuned": {"panel_width": block, "zgemm_tile": "64x64" if cfg == 1 else "64x32", "graph_ms": round(tuned_ms, 2),
                      "tuning_s": round(tune_s, 2)},
            "singular": int((sing >= 0).sum()), "max_backward_error": worst,
            "rcond_range": [float(rc.min()), float(rc.max())], "cpu_baseline": cpu}


def grad_leg(ws, sizes=(256, 1024), cpu_sample=False):
    """Gradients of L = ||C exp(tA) B - Y||_F^2 (t = 1, ||A||_1 = 20, B n x 4,
    C 8 x n, Y 8 x 4) for real and complex A: forward Expmv + the adjoint
    (dL/dA, dL/dB), with the error against the 2n x 2n block exponential
    (computed by this library's Expm, on the device)."""
    import torch

    from csharp_linearalgebra_amd import workloads as W

    out = []
    for cplx in (False, True):
        for n in sizes:
            a, b, c, y = W.adjoint_problem(n, cplx, target=20.0, seed=0)
            dA, dB = to_device(a), to_device(b)
            f, _ = ws.expmv(a, b, 1.0)
            g = 2.0 * c.conj().T @ (c @ f - y)
            dG = to_device(g)
            ga = empty_device(n, n, cplx)
            gb = empty_device(n, 4, cplx)
            dF = empty_device(n, 4, cplx)
            torch.cuda.synchronize()
            ws.expmv_backward(dA, dB, 1.0, dG, out_a=ga, out_b=gb)
            ws.synchronize()
            t0 = time.perf_counter()
            ws.expmv(dA, dB, 1.0, out=dF)
            ws.synchronize()
            t_fwd = time.perf_counter() - t0
            t0 = time.perf_counter()
            ws.expmv_backward(dA, dB, 1.0, dG, out_a=ga, out_b=gb)
            ws.synchronize()
            t_bwd = time.perf_counter() - t0
            # Block-exponential reference of dL/dA = L(A^H, G B^H).
            blk = np.zeros((2 * n, 2 * n), dtype=np.complex128 if cplx else np.float64)
            blk[:n, :n] = a.conj().T
            blk[n:, n:] = a.conj().T
            blk[:n, n:] = g @ b.conj().T
            e, _ = ws.expm(blk)
            ra = e[:n, n:]
            got = ga.owner.cpu().numpy().T
            err = float(np.abs(got - ra).max() / np.abs(ra).max())
            row = {"n": n, "complex": cplx, "forward_ms": round(t_fwd * 1e3, 2),
                   "backward_ms": round(t_bwd * 1e3, 2), "grad_a_rel_error_vs_block_expm": err}
            if cpu_sample and n <= 256:
                import oracle as O

                # The host's way to this gradient: the oracle's Expm of the
                # 2n x 2n block matrix (the reference has no adjoint).
                t0 = time.perf_counter()
                O.expm(blk)
                cs = time.perf_counter() - t0
                row["cpu_baseline"] = {"value": round(cs * 1e3, 1), "unit": "ms", "cores": O.threads(),
                                       "kind": "port",
                                       "sample": "dL/dA by the oracle's Expm of the %dx%d block matrix" % (2 * n, 2 * n)}
            out.append(row)
    return out


def expm_leg(ws, n=1000, steps=3):
    """The benchmark step alone (complex Expm, n = 1000, ||A||_1 = 50), one
    warm-up and `steps` timed, for rocprofv3."""
    import torch

    from csharp_linearalgebra_amd import workloads as W

    a, _ = W.expm_operator(n, 50.0, 8, seed=0)
    dA = to_device(a)
    dE = empty_device(n, n, True)
    torch.cuda.synchronize()
    ws.expm(dA, out=dE)
    ws.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ws.expm(dA, out=dE)
    ws.synchronize()
    return {"expm_ms": round((time.perf_counter() - t0) / steps * 1e3, 3)}


def expmv_leg(ws, n=1000, reps=3):
    """The Expmv of the report alone (complex, n = 1000, B n x 8, t = 0.5),
    one warm-up and `reps` timed, for rocprofv3."""
    import torch

    from csharp_linearalgebra_amd import workloads as W

    a, b = W.expm_operator(n, 50.0, 8, seed=0)
    dA, dB = to_device(a), to_device(b)
    dF = empty_device(n, 8, True)
    torch.cuda.synchronize()
    ws.expmv(dA, dB, 0.5, out=dF)
    ws.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ws.expmv(dA, dB, 0.5, out=dF)
    ws.synchronize()
    return {"expmv_ms": round((time.perf_counter() - t0) / reps * 1e3, 3)}


def kernels_leg(ws, n=8192):
    """The memory-bound parts on their own, for rocprofv3: the 1-norm and
    the infinity-norm of an n x n FP64 matrix (8 n^2 bytes each), the unit
    lower triangular solve with one right-hand side (the triangle, 4 n^2
    bytes), and an FP64 LU of order 4096 (its panels and trailing updates).
    Returns the solve's HIP-event time on the library stream."""
    import ctypes as C

    import torch

    lib = ws._lib
    a = torch.rand((n, n), dtype=torch.float64, device="cuda") + torch.eye(n, dtype=torch.float64, device="cuda") * n
    x = torch.rand((1, n), dtype=torch.float64, device="cuda")
    out = C.c_double()
    torch.cuda.synchronize()
    for kind in (b"1", b"I"):
        for _ in range(10):
            ws.check(lib.tns_dnorm(ws.handle, C.c_char(kind), n, n, a.data_ptr(), n, C.byref(out)))
    trsv = event_time(ws, lambda: ws.check(lib.tns_dtrsm(ws.handle, 0, n, 1, a.data_ptr(), n, x.data_ptr(), n)), 5)
    m = 4096
    f = torch.rand((m, m), dtype=torch.float64, device="cuda")
    piv = torch.zeros(m, dtype=torch.int32, device="cuda")
    from csharp_linearalgebra_amd import _lib as L

    info = L.LuInfo()
    torch.cuda.synchronize()
    ws.check(lib.tns_dgetrf(ws.handle, m, m, f.data_ptr(), m, 0, piv.data_ptr(), C.byref(info)))
    ws.synchronize()
    return {"trsv_n%d_us" % n: round(trsv * 1e6, 1), "trsv_gbs": round(4.0 * n * n / trsv / 1e9, 1)}


def lu_device(ws, la, lb):
    """Config 1 with every operand in device memory: getrf, the 16-RHS getrs
    and gecon through the C ABI (best of three)."""
    import ctypes as C

    import torch

    from csharp_linearalgebra_amd import _lib as L

    lib = ws._lib
    n, nrhs = la.shape[0], lb.shape[1]
    at = torch.from_numpy(np.ascontiguousarray(la.T)).cuda()
    bt = torch.from_numpy(np.ascontiguousarray(lb.T)).cuda()
    f, x = torch.empty_like(at), torch.empty_like(bt)
    piv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = L.LuInfo()
    anorm = float(np.abs(la).sum(0).max())
    best = None
    for _ in range(4):
        f.copy_(at)
        x.copy_(bt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ws.check(lib.tns_dgetrf(ws.handle, n, n, f.data_ptr(), n, 0, piv.data_ptr(), C.byref(info)))
        t1 = time.perf_counter()
        ws.check(lib.tns_dgetrs(ws.handle, 0, n, nrhs, f.data_ptr(), n, piv.data_ptr(), -1, x.data_ptr(), n))
        ws.synchronize()
        t2 = time.perf_counter()
        rc = C.c_double()
        ws.check(lib.tns_dgecon(ws.handle, n, f.data_ptr(), n, piv.data_ptr(), -1, anorm, 2, C.byref(rc)))
        t3 = time.perf_counter()
        row = (t1 - t0, t2 - t1, t3 - t2)
        best = row if best is None or sum(row) < sum(best) else best
    return {"getrf_ms": round(best[0] * 1e3, 3), "getrs16_ms": round(best[1] * 1e3, 3),
            "gecon_ms": round(best[2] * 1e3, 3), "total_ms": round(sum(best) * 1e3, 3)}


def lu_leg(ws, n=8192):
    """For rocprofv3: one FP64 LU of order n (device-resident, look-ahead)
    and three FP64 GEMMs n x n x n, on uniform(-1, 1) matrices."""
    import ctypes as C

    import torch

    from csharp_linearalgebra_amd import _lib as L

    lib = ws._lib
    g = torch.Generator(device="cuda").manual_seed(0)
    a = torch.rand((n, n), dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    f = a.clone()
    piv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = L.LuInfo()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ws.check(lib.tns_dgetrf(ws.handle, n, n, f.data_ptr(), n, 0, piv.data_ptr(), C.byref(info)))
    ws.synchronize()
    t_lu = time.perf_counter() - t0
    c = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        ws.check(lib.tns_dgemm(ws.handle, n, n, n, 1.0, a.data_ptr(), n, f.data_ptr(), n, 0.0, c.data_ptr(), n))
    ws.synchronize()
    return {"getrf_n%d_ms" % n: round(t_lu * 1e3, 2)}


def load_traffic():
    path = os.path.join(ROOT, "profiles", "zgemm_n1000_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--roofline-only", action="store_true", help="run only the dominant kernel (for rocprofv3)")
    ap.add_argument("--leg", choices=["sweep", "mixed", "grad", "kernels", "expm", "expmv", "lu"],
                    help="run one extra leg alone and print its JSON (for rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this backend has no multi-GPU path (run with --gpus 1).", file=sys.stderr)
        sys.exit(2)
    if args.roofline_only:
        roofline_only()
        return
    if args.leg:
        import csharp_linearalgebra_amd as T

        ws = T.Workspace(0)
        if args.leg == "sweep":
            out = sweep_leg(ws)
        elif args.leg == "mixed":
            out = mixed_leg(ws, 8192, (1e2,))
        elif args.leg == "kernels":
            out = kernels_leg(ws)
        elif args.leg == "expm":
            out = expm_leg(ws)
        elif args.leg == "expmv":
            out = expmv_leg(ws)
        elif args.leg == "lu":
            out = lu_leg(ws)
        else:
            out = grad_leg(ws)
        print(json.dumps(out), flush=True)
        return

    import torch

    import csharp_linearalgebra_amd as T
    from csharp_linearalgebra_amd import workloads as W

    ws = T.Workspace(0)
    n = 1000
    a, b = W.expm_operator(n, 50.0, 8, seed=0)
    dA = to_device(a)
    dE = empty_device(n, n, True)
    dB = to_device(b)
    dF = empty_device(n, 8, True)
    torch.cuda.synchronize()

    # ---- the step: complex Expm, n = 1000 ----------------------------------------
    for _ in range(args.warmup):
        _, diag = ws.expm(dA, out=dE)
    ws.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, diag = ws.expm(dA, out=dE)
    ws.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    step_s = (t1 - t0) / args.steps
    flops = W.expm_flops(n, diag.squarings, True)
    gflops = flops / step_s / 1e9

    # ---- the step's split: three more steps with HIP events between phases ------
    ph = T.ExpmPhaseTimings()
    for _ in range(3):
        ws.expm(dA, out=dE, phases=ph)
    phases = ph.per_call()
    families = {"gemm_products": round(phases["powers"] + phases["pade"] + phases["squaring"], 3),
                "norm_estimates": phases["select"], "lu": phases["lu"], "solve": phases["solve"],
                "total": phases["total"]}
    # The LU's own split at the denominator's order (the sequential loop:
    # its phase clock turns the look-ahead off).
    lph = T.LuPhaseTimings()
    dL = empty_device(n, n, True)
    for _ in range(3):
        dL.owner.copy_(dA.owner)
        torch.cuda.synchronize()
        ws.factor_lu_device(dL, timings=lph)
    lu_split = {k: round(v / 3, 3) for k, v in (("panel", lph.panel), ("swaps", lph.swaps),
                                                ("trsm", lph.triangular), ("gemm", lph.gemm),
                                                ("total", lph.total))}

    # ---- dominant kernel: 4M complex GEMM, n = 1000 ---------------------------------
    dC = empty_device(n, n, True)
    g_s = launch_time(ws, lambda: ws.multiply(dA, dA, dC), 20)
    zg_tflops = 8.0 * n ** 3 / g_s / 1e12

    # ---- exponential action exp(tA)B, B n x 8, t = 0.5 ---------------------------
    _, vd = ws.expmv(dA, dB, 0.5, out=dF)
    ws.synchronize()
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        _, vd = ws.expmv(dA, dB, 0.5, out=dF)
    ws.synchronize()
    expmv_s = (time.perf_counter() - t0) / reps
    expmv_flops = vd.applications * 8.0 * n * n * 8

    # ---- FP64 LU + solve + rcond, n = 1024 ---------------------------------------
    la, lb = W.lu_system(1024, 16, seed=0)
    lu = ws.factor_lu(la)
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        lu = ws.factor_lu(la)
    lu_s = (time.perf_counter() - t0) / reps
    x = lu.solve(lb)
    rc = lu.reciprocal_condition()
    t0 = time.perf_counter()
    for _ in range(reps):
        x = lu.solve(lb)
        rc = lu.reciprocal_condition()
    solve_s = (time.perf_counter() - t0) / reps
    lu_bwd = float(np.abs(la @ x - lb).max() / (np.abs(la).sum(1).max() * np.abs(x).max()))
    dev = lu_device(ws, la, lb)

    # ---- FP64 GEMM, n = 8192 -----------------------------------------------------
    ng = 8192
    g = np.random.default_rng(1)
    dX = to_device(g.uniform(-1, 1, (ng, ng)))
    dY = to_device(g.uniform(-1, 1, (ng, ng)))
    dZ = empty_device(ng, ng, False)
    torch.cuda.synchronize()
    dg_s = event_time(ws, lambda: ws.multiply(dX, dY, dZ), 5)
    dg_tflops = 2.0 * ng ** 3 / dg_s / 1e12

    result = {
        "metric": "GFLOP/s & % FP64-MFMA peak; MP-LU+IR vs FP64 LU speedup @1e-15 bwd err",
        "value": round(gflops, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128 (f64 MFMA)",
        "data": "synthetic, seeded (numpy PCG64 seed 0)",
        "config": {"workload": "complex128 Expm, n=1000, A=-D+iW, ||A||_1=50 (degree 13, s=%d)" % diag.squarings,
                   "n": n, "degree": diag.degree, "squarings": diag.squarings},
        "roofline": {"bound": "mfma", "kernel": "zgemm_kernel (4M, n=1000)", "achieved": round(zg_tflops, 2),
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(zg_tflops / FP64_PEAK_TFLOPS, 4),
                     "traffic": load_traffic(), "avg_launch_us": round(g_s * 1e6, 1),
                     "algorithmic_flops_per_launch": 8.0 * n ** 3},
        "expmv": {"workload": "complex128 exp(tA)B, n=1000, B 1000x8, t=0.5", "ms": round(expmv_s * 1e3, 3),
                  "degree": vd.degree, "scaling": vd.scaling, "applications": vd.applications,
                  "gflops": round(expmv_flops / expmv_s / 1e9, 2)},
        "step_phases_ms": phases,
        "step_families_ms": families,
        "lu_phase_split_n1000_complex_ms": lu_split,
        "lu_n1024": {"host_operands": {"getrf_ms": round(lu_s * 1e3, 3),
                                       "getrf_gflops": round(W.getrf_flops(1024) / lu_s / 1e9, 1),
                                       "solve16_plus_rcond_ms": round(solve_s * 1e3, 3)},
                     "device_resident": dev, "rcond": rc, "backward_error": lu_bwd,
                     "identity_pivots": bool(np.array_equal(lu.pivots, np.arange(1024)))},
        "dgemm_n8192": {"tflops": round(dg_tflops, 2), "frac": round(dg_tflops / FP64_PEAK_TFLOPS, 4)},
    }

    # ---- mixed-precision LU + FP64 refinement vs FP64 LU, n = 4096 and 8192 -------
    result["mixed_lu"] = {str(nm): mixed_leg(ws, nm, [1e2, 1e6, 1e10]) for nm in (4096, 8192)}
    # The metric's second figure: MP-LU + IR against the FP64 LU, n = 8192,
    # at the conditioning where refinement converges (kappa = 1e2).
    row = result["mixed_lu"]["8192"][0]
    result["mixed_speedup_n8192"] = {"kappa": row["kappa"], "fp64_ms": row["fp64_ms"],
                                     "fp32": row["fp32"]["speedup"], "bf16": row["bf16"]["speedup"],
                                     "backward_error_fp32": row["fp32"]["backward_error"],
                                     "fp32_partial_pivoting_ir": row["fp32_partial_pivoting_ir"]["speedup"]}

    # ---- batched frequency sweep ------------------------------------------------
    result["batched_sweep"] = sweep_leg(ws, cpu_sample=not args.no_cpu)

    # ---- gradients through exp(tA)B ---------------------------------------------
    result["expmv_gradients"] = grad_leg(ws, cpu_sample=not args.no_cpu)

    if not args.no_cpu:
        import oracle as O

        # Config 1 on the host: LU, 16-RHS solve and rcond of the same system.
        reps, t0 = 0, time.perf_counter()
        while True:
            f = O.getrf(la)
            O.getrs(f, lb)
            O.gecon(f)
            reps += 1
            cs = time.perf_counter() - t0
            if cs >= 3.0:
                break
        result["lu_n1024"]["cpu_baseline"] = {
            "value": round(cs / reps * 1e3, 2), "unit": "ms per LU+solve16+rcond", "cores": O.threads(),
            "kind": "port", "gflops": round(W.getrf_flops(1024) * reps / cs / 1e9, 1),
            "sample": "%d runs of the same n=1024 system, %.1f s" % (reps, cs)}
        # The FP64 LU the mixed solves are measured against, at n = 4096, on
        # the host (one factorization of a matrix of the same kind).
        g4 = np.random.default_rng(0)
        m4 = g4.uniform(-1, 1, (4096, 4096))
        t0 = time.perf_counter()
        O.getrf(m4)
        cs = time.perf_counter() - t0
        result["mixed_lu"]["cpu_baseline"] = {
            "value": round(W.getrf_flops(4096) / cs / 1e9, 1), "unit": "GFLOP/s (FP64 LU)", "cores": O.threads(),
            "kind": "port", "ms": round(cs * 1e3, 1), "sample": "one FP64 LU of a 4096 x 4096 matrix"}

        # A bounded sample: whole exponentials of the same operator until
        # about 10 s of host time (at most 30 s) have passed.
        reps, cflops, t0 = 0, 0.0, time.perf_counter()
        while True:
            _, cdiag = O.expm(a)
            reps += 1
            cflops += W.expm_flops(n, cdiag[1], True)
            cpu_s = time.perf_counter() - t0
            if cpu_s >= 10.0 or cpu_s * (reps + 1) / reps > 30.0:
                break
        result["cpu_baseline"] = {"value": round(cflops / cpu_s / 1e9, 2), "unit": "GFLOP/s",
                                  "cores": O.threads(), "kind": "port",
                                  "sample": "%d complex128 Expm of the same n=1000 operator (degree %d, s=%d; "
                                            "C restatement of Tensile, OpenMP AVX2 GEMM), %.1f s"
                                            % (reps, cdiag[0], cdiag[1], cpu_s)}
    print(json.dumps(result))
    ws.dispose()


if __name__ == "__main__":
    main()
