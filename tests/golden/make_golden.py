#!/usr/bin/env python3
"""Writes dense_path_golden.npz: independent reference values for the
oracle's tests, from LAPACK (through scipy) and from mpmath at 50 digits.

  * LAPACK dgetrf / zgetrf pivots and factors of seeded matrices: the same
    idamax / izamax rule (ties to the lowest index, |Re|+|Im| for complex)
    Tensile's LU follows (Pivoting.cs:16-23, ElementKernels.cs:61-69).
  * LAPACK dgecon / zgecon reciprocal condition numbers (same estimator
    family; Tensile's is reproduced only in order of magnitude).
  * exp(A) of small real and complex matrices by a 50-digit Taylor series
    with scaling and squaring in mpmath: a reference three orders beyond
    double precision, independent of any library's expm.
  * exp(tA) B for the same matrices.

Run from the repository root: python tests/golden/make_golden.py
"""
import os

import mpmath as mp
import numpy as np
import scipy.linalg as sl

HERE = os.path.dirname(os.path.abspath(__file__))


def mp_expm(a):
    mp.mp.dps = 50
    n = a.shape[0]
    A = mp.matrix(n, n)
    for i in range(n):
        for j in range(n):
            A[i, j] = mp.mpc(complex(a[i, j]))
    norm = max(sum(abs(A[i, j]) for i in range(n)) for j in range(n))
    s = 0
    while norm > mp.mpf(1) / 64:
        norm /= 2
        s += 1
    B = A / (mp.mpf(2) ** s)
    E = mp.eye(n)
    term = mp.eye(n)
    for k in range(1, 80):
        term = term * B / k
        E = E + term
    for _ in range(s):
        E = E * E
    out = np.zeros((n, n), dtype=complex)
    for i in range(n):
        for j in range(n):
            out[i, j] = complex(E[i, j])
    return out


def main():
    rng = np.random.default_rng(20260915)
    data = {}
    for idx, (m, n) in enumerate([(7, 7), (40, 40), (65, 65), (50, 37), (37, 50)]):
        a = rng.uniform(-1, 1, (m, n))
        lu, piv, _ = sl.lapack.dgetrf(a)
        data[f"dlu{idx}_a"], data[f"dlu{idx}_lu"], data[f"dlu{idx}_piv"] = a, lu, piv
        z = rng.uniform(-1, 1, (m, n)) + 1j * rng.uniform(-1, 1, (m, n))
        zlu, zpiv, _ = sl.lapack.zgetrf(z)
        data[f"zlu{idx}_a"], data[f"zlu{idx}_lu"], data[f"zlu{idx}_piv"] = z, zlu, zpiv
    for idx, n in enumerate([10, 60]):
        a = rng.uniform(-1, 1, (n, n))
        lu, piv, _ = sl.lapack.dgetrf(a)
        rc, _ = sl.lapack.dgecon(lu, np.abs(a).sum(0).max(), norm="1")
        data[f"dgecon{idx}_a"], data[f"dgecon{idx}_rc"] = a, rc
        data[f"dgecon{idx}_exact"] = 1.0 / (np.linalg.norm(a, 1) * np.linalg.norm(np.linalg.inv(a), 1))
    cases = [("e0", 6, 0.05, False), ("e1", 8, 0.9, False), ("e2", 8, 6.0, False), ("e3", 10, 40.0, False),
             ("e4", 6, 0.3, True), ("e5", 9, 3.0, True), ("e6", 12, 25.0, True)]
    for name, n, norm, cplx in cases:
        a = rng.standard_normal((n, n))
        if cplx:
            a = a + 1j * rng.standard_normal((n, n))
        a *= norm / np.abs(a).sum(0).max()
        b = rng.standard_normal((n, 3)) + (1j * rng.standard_normal((n, 3)) if cplx else 0)
        data[f"{name}_a"] = a
        e = mp_expm(a)
        data[f"{name}_exp"] = e if cplx else e.real
        et = mp_expm(0.5 * a)
        data[f"{name}_b"] = b
        data[f"{name}_expmv"] = (et @ b) if cplx else (et.real @ b)
    np.savez_compressed(os.path.join(HERE, "dense_path_golden.npz"), **data)
    print("wrote", len(data), "arrays")


if __name__ == "__main__":
    main()
