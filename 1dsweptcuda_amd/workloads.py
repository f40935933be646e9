"""Seeded synthetic workloads of the benchmark configurations.

Batched uneven solves: `count` independent periodic problems, a seeded 50/50
mix of heat and Kuramoto-Sivashinsky, nX = floor(2^U[10,16]), Fourier-mode
initial conditions, dt at 0.8x the linear stability limit:
  heat  dx = 1/nX (L = 1), alpha = 8.418e-5: FTCS needs Fo = alpha*dt/dx^2 <= 1/2,
        dt = 0.8 * 0.5 * dx^2 / alpha;
  KS    dx = 0.5: the RK2-midpoint factor 1 + z + z^2/2 of the most unstable
        discrete mode, z = dt*(4/dx^2 - 16/dx^4), needs z >= -2,
        dt = 0.8 * 2 / (16/dx^4 - 4/dx^2).
"""
from __future__ import annotations

import numpy as np

ALPHA = 8.418e-5
HEAT, KS = 0, 1


def batch_problems(count: int = 4096, seed: int = 2024, lo: float = 10.0, hi: float = 16.0):
    rng = np.random.default_rng(seed)
    eq = np.where(rng.random(count) < 0.5, HEAT, KS).astype(np.int32)
    nx = np.floor(2.0 ** rng.uniform(lo, hi, count)).astype(np.int64)
    dx = np.where(eq == HEAT, 1.0 / nx, 0.5)
    dt = np.where(eq == HEAT, 0.8 * 0.5 * dx * dx / ALPHA,
                  0.8 * 2.0 / (16.0 / dx ** 4 - 4.0 / dx ** 2))
    return eq, nx, dt, dx


def fourier_ic(n: int, L: float, seed: int, K: int = 16, offset: float = 0.0) -> np.ndarray:
    """u(x) = offset + sum_{k<=K} (a_k cos + b_k sin)(2 pi k x / L) / k, a, b ~ N(0,1)."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((K, 2))
    x = np.arange(n) * (L / n)
    u = np.full(n, offset)
    for k in range(1, K + 1):
        th = 2 * np.pi * k * x / L
        u += (a[k - 1, 0] * np.cos(th) + a[k - 1, 1] * np.sin(th)) / k
    return u


def batch_ics(eq, nx, dx, seed: int = 7):
    out = []
    for i, (e, n, d) in enumerate(zip(eq, nx, dx)):
        L = 1.0 if e == HEAT else n * d
        out.append(fourier_ic(int(n), L, seed + i, offset=1.0 if e == HEAT else 0.0))
    return out
