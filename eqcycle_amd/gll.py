"""Gauss-Legendre-Lobatto nodes, weights and derivative matrix for NGLL points.

Replaces the table lookup ``GetGLL(ngll)`` (src/GetGLL.jl:14-24), which reads
``src/gll_xwh/gll_NN.tab``.  The values here are computed from the closed forms,
so no table ships with the package:

* nodes: -1, +1 and the roots of P'_{N}(x) (N = NGLL-1), found by Newton on
  the Legendre recurrence;
* weights: w_j = 2 / (N (N+1) P_N(x_j)^2);
* derivative: D[p, i] = h'_i(x_p) = P_N(x_p) / (P_N(x_i) (x_p - x_i)) for p != i,
  D[0, 0] = -N(N+1)/4, D[N, N] = +N(N+1)/4, zero otherwise on the diagonal.

The reference stores ``H[i, j] = h'_i(x_j)``, i.e. ``H = D^T``.
"""
from __future__ import annotations

import numpy as np


def _legendre(n: int, x: np.ndarray):
    """P_n(x) and P_n'(x) by the three-term recurrence."""
    p0 = np.ones_like(x)
    if n == 0:
        return p0, np.zeros_like(x)
    p1 = x.copy()
    for k in range(2, n + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    # derivative from (1-x^2) P_n' = n (P_{n-1} - x P_n), valid off the endpoints
    with np.errstate(divide="ignore", invalid="ignore"):
        dp = n * (p0 - x * p1) / (1.0 - x * x)
    return p1, dp


def gll(ngll: int = 5):
    """Return (xgll, wgll, H) with H[i, j] = h'_i(xgll[j]) (the reference layout)."""
    if ngll < 2:
        raise ValueError("ngll must be >= 2")
    n = ngll - 1
    # Chebyshev-Gauss-Lobatto start, Newton on (1-x^2) P_n'(x) = 0 for the interior
    x = -np.cos(np.pi * np.arange(ngll) / n)
    for _ in range(100):
        xi = x[1:-1]
        p, dp = _legendre(n, xi)
        # q(x) = P_n'(x); q'(x) from the Legendre ODE: (1-x^2) P'' = 2x P' - n(n+1) P
        d2p = (2 * xi * dp - n * (n + 1) * p) / (1.0 - xi * xi)
        step = dp / d2p
        x[1:-1] = xi - step
        if np.max(np.abs(step)) < 1e-17:
            break
    x[0], x[-1] = -1.0, 1.0
    # symmetrise so x[k] == -x[n-k] exactly
    x = 0.5 * (x - x[::-1])
    if ngll % 2 == 1:
        x[n // 2] = 0.0
    pn, _ = _legendre(n, x)
    w = 2.0 / (n * (n + 1) * pn * pn)
    D = np.zeros((ngll, ngll))
    for p in range(ngll):
        for i in range(ngll):
            if p != i:
                D[p, i] = pn[p] / (pn[i] * (x[p] - x[i]))
    D[0, 0] = -n * (n + 1) / 4.0
    D[n, n] = n * (n + 1) / 4.0
    H = D.T.copy()
    return x, w, H
