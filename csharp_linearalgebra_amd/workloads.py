"""Seeded synthetic inputs for the benchmark workloads and their parity tests.

Each generator builds one of the dense-path workloads from numpy's PCG64
generator with an explicit seed, so the device and the CPU oracle see the
same arrays. Shapes and distributions:

  lu_system         n=1024 real: A i.i.d. U(-1,1) + n*I, B n x 16 U(-1,1)
  expm_operator     complex n=1000: A = -D + i*W, D = diag(U(0.1,10)) plus a
                    symmetric U(-0.05,0.05) coupling, W symmetric U(-1,1)
                    scaled so ||A||_1 = target (50); B n x 8 complex U(-1,1)
  conditioned       real n: A = U diag(sigma) V^T with U, V products of 64
                    Householder reflectors, sigma log-spaced to condition
                    kappa; B n x 32 Gaussian
  frequency_sweep   count complex matrices Z(w) = R + i w L of uneven order,
                    R = diag U(0.01,1), L = G G^T / n + I, w log-spaced
  adjoint_problem   the loss data of the exp(tA)B gradient workload
"""
from __future__ import annotations

import numpy as np


def lu_system(n: int = 1024, nrhs: int = 16, seed: int = 0):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1.0, 1.0, (n, n)) + n * np.eye(n)
    b = rng.uniform(-1.0, 1.0, (n, nrhs))
    return np.asfortranarray(a), np.asfortranarray(b)


def _one_norm(a):
    return np.abs(a).sum(axis=0).max()


def expm_operator(n: int = 1000, target: float = 50.0, nrhs: int = 8, seed: int = 0):
    """A = -D + iW with ||A||_1 = target (to rounding), and B."""
    rng = np.random.default_rng(seed)
    d = np.diag(rng.uniform(0.1, 10.0, n))
    c = rng.uniform(-0.05, 0.05, (n, n))
    c = np.triu(c, 1)
    d = d + c + c.T
    w = rng.uniform(-1.0, 1.0, (n, n))
    w = np.triu(w) + np.triu(w, 1).T
    lo, hi = 0.0, 1.0
    while _one_norm(-d + 1j * hi * w) < target:
        hi *= 2.0
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        if _one_norm(-d + 1j * mid * w) < target:
            lo = mid
        else:
            hi = mid
    a = -d + 1j * (0.5 * (lo + hi)) * w
    b = rng.uniform(-1.0, 1.0, (n, nrhs)) + 1j * rng.uniform(-1.0, 1.0, (n, nrhs))
    return np.asfortranarray(a), np.asfortranarray(b)


def _householder_product(n, count, rng):
    q = np.eye(n)
    for _ in range(count):
        v = rng.standard_normal(n)
        v /= np.linalg.norm(v)
        q = q - 2.0 * np.outer(v, v @ q)
    return q


def conditioned(n: int, kappa: float, nrhs: int = 32, seed: int = 0):
    rng = np.random.default_rng(seed)
    u = _householder_product(n, 64, rng)
    v = _householder_product(n, 64, rng)
    sigma = np.logspace(0.0, -np.log10(kappa), n)
    a = (u * sigma) @ v.T
    b = rng.standard_normal((n, nrhs))
    return np.asfortranarray(a), np.asfortranarray(b)


def conditioned_device(n: int, kappa: float, nrhs: int = 32, seed: int = 0):
    """conditioned() built on the device with torch: the same seeded draws,
    the reflectors applied to diag(sigma) as rank-one updates (O(n^2) each),
    so n = 8192 takes milliseconds. Returns (A, B) as torch float64 matrices
    on cuda:0 (logical row/column order; transpose for column-major storage)."""
    import torch

    rng = np.random.default_rng(seed)
    def draws():
        out = []
        for _ in range(64):
            v = rng.standard_normal(n)
            out.append(v / np.linalg.norm(v))
        return torch.from_numpy(np.stack(out)).cuda()
    hu = draws()
    hv = draws()
    b = torch.from_numpy(rng.standard_normal((n, nrhs))).cuda()
    sigma = torch.from_numpy(np.logspace(0.0, -np.log10(kappa), n)).cuda()
    a = torch.diag(sigma)
    for k in range(64):  # A <- H_k A, k = 1..64: U = H_64 ... H_1
        v = hu[k]
        a -= 2.0 * torch.outer(v, v @ a)
    for k in range(64):  # A <- A G_k: A = U S G_1 ... G_64 = U S V^T
        g = hv[k]
        a -= 2.0 * torch.outer(a @ g, g)
    return a, b


def frequency_sweep(count: int = 512, nmin: int = 64, nmax: int = 1024, seed: int = 0):
    rng = np.random.default_rng(seed)
    orders = rng.integers(nmin, nmax + 1, count)
    omegas = np.logspace(1.0, 6.0, count)
    mats = []
    for n, w in zip(orders, omegas):
        r = np.diag(rng.uniform(0.01, 1.0, n))
        g = rng.standard_normal((n, n))
        l = g @ g.T / n + np.eye(n)
        mats.append(np.asfortranarray(r + 1j * w * l))
    return mats


def frequency_sweep_device(count: int = 512, nmin: int = 64, nmax: int = 1024, nrhs: int = 4, seed: int = 0):
    """frequency_sweep() built on the device for the benchmark: the orders,
    frequencies and R diagonals from numpy's seeded generator, G and the
    right-hand sides (complex uniform(-1,1)) from a seeded torch generator
    on cuda:0. Returns (mats, rhs): lists of complex128 torch tensors in
    column-major storage (each tensor holds the transpose, shape (cols,
    rows))."""
    import torch

    rng = np.random.default_rng(seed)
    orders = rng.integers(nmin, nmax + 1, count)
    omegas = np.logspace(1.0, 6.0, count)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    mats, rhs = [], []
    for n, w in zip(orders, omegas):
        n = int(n)
        r = torch.from_numpy(rng.uniform(0.01, 1.0, n)).cuda()
        g = torch.randn((n, n), generator=gen, device="cuda", dtype=torch.float64)
        l = g @ g.T / n + torch.eye(n, device="cuda", dtype=torch.float64)
        z = torch.complex(torch.diag(r), w * l)  # symmetric: its transpose is itself
        mats.append(z.contiguous())
        re = torch.rand((nrhs, n), generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
        im = torch.rand((nrhs, n), generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
        rhs.append(torch.complex(re, im).contiguous())
    return mats, rhs


def adjoint_problem(n: int, complex_: bool, target: float = 20.0, seed: int = 0):
    rng = np.random.default_rng(seed)
    if complex_:
        a, _ = expm_operator(n, target, 1, seed)
    else:
        a, _ = expm_operator(n, target, 1, seed)
        a = np.asfortranarray(a.real * (target / _one_norm(a.real)))
    dt = np.complex128 if complex_ else np.float64
    def g(*shape):
        x = rng.standard_normal(shape)
        if complex_:
            x = x + 1j * rng.standard_normal(shape)
        return np.asfortranarray(x.astype(dt))
    return a, g(n, 4), g(8, n), g(8, 4)


def expm_flops(n: int, squarings: int, complex_: bool = True) -> float:
    """Algorithmic flops of a degree-13 Expm with s squarings: (6 + s)
    products, the LU of the Pade denominator and its n-column solve."""
    mul = 8.0 if complex_ else 2.0
    lu = (8.0 / 3.0 if complex_ else 2.0 / 3.0) * n ** 3
    solve = (8.0 if complex_ else 2.0) * n ** 3
    return (6 + squarings) * mul * n ** 3 + lu + solve


def sweep_flops(orders, nrhs: int = 4) -> float:
    """Complex LU (8/3 n^3) and the solve (8 n^2 nrhs) of every system."""
    return float(sum((8.0 / 3.0) * n ** 3 + 8.0 * n * n * nrhs for n in orders))


def getrf_flops(n: int) -> float:
    return (2.0 / 3.0) * n ** 3 - n ** 2 / 2.0 - n / 6.0
