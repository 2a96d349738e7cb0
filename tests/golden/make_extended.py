#!/usr/bin/env python3
"""Writes expm_n1000_extended.npz: extended-precision references for the
benchmark's complex operator of order 1000 (A = -D + iW, ||A||_1 = 50, and
B 1000 x 8 from csharp_linearalgebra_amd.workloads.expm_operator(seed=0)):

  * exp(A) V for V = the first four columns of B,
  * exp(A/2) B (the Expmv of the benchmark, t = 0.5),

both by the oracle's x87 long-double Taylor action (or_zexpmv_taylor_ld:
tA/s with ||tA/s||_1 <= 1, Taylor sums to 1e-22; checked against the 50-digit
mpmath references of dense_path_golden.npz to 7e-18). A checksum of A and B
is stored so a change of the generator is caught instead of silently
comparing against another matrix.

Run from the repository root: python tests/golden/make_extended.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from csharp_linearalgebra_amd import workloads as W  # noqa: E402


def main():
    a, b = W.expm_operator(1000, 50.0, 8, seed=0)
    t0 = time.time()
    ev = O.expmv_taylor_ld(a, b[:, :4], 1.0)
    t1 = time.time()
    eb = O.expmv_taylor_ld(a, b, 0.5)
    t2 = time.time()
    print(f"exp(A)V {t1 - t0:.1f} s, exp(A/2)B {t2 - t1:.1f} s")
    np.savez_compressed(os.path.join(HERE, "expm_n1000_extended.npz"), expm_v=ev, expmv_b=eb,
                        checksum=np.array([a.sum(), (a * a).sum(), b.sum(), np.abs(a).sum(0).max()]))


if __name__ == "__main__":
    main()
