"""Gauss-Jacobi type quadratures on [-1, 1] by the Golub-Welsch eigenvalue
method (the algorithm of frfs/quadratures.py:6-55, after W. Gautschi,
"Orthogonal Polynomials: Computation and Approximation").

zwgj  : Gauss-Jacobi(a, b)            (frfs/quadratures.py:40-43)
zwglj : Gauss-Lobatto-Jacobi(a, b)    (frfs/quadratures.py:45-55)
zwgrjm/zwgrjp : Gauss-Radau-Jacobi with the node -1 / +1, the rules the
    multi-species mesh names (frfs/solvers/dgfsbi/velocitymesh.py:127-132)
    but never defines.
"""
from math import gamma as _gamma

import numpy as np


def _jacobi_recurrence(n, a, b):
    """Monic Jacobi recurrence: alpha_k (k<n) and beta_k (k<n), beta_0 = mu_0."""
    al = np.zeros(n)
    be = np.zeros(n)
    s = a + b
    for k in range(n):
        d = 2.0 * k + s
        al[k] = (b - a) / (2.0 + s) if k == 0 else (b * b - a * a) / (d * (d + 2.0))
    be[0] = 2.0 ** (s + 1) * _gamma(a + 1) * _gamma(b + 1) / _gamma(s + 2)
    if n > 1:
        be[1] = 4.0 * (a + 1) * (b + 1) / ((s + 3.0) * (s + 2.0) ** 2)
    for k in range(2, n):
        d = 2.0 * k + s
        be[k] = 4.0 * k * (k + a) * (k + b) * (k + s) / ((d * d - 1.0) * d * d)
    return al, be


def _golub_welsch(al, be):
    n = len(al)
    off = np.sqrt(be[1:n])
    T = np.diag(al) + np.diag(off, 1) + np.diag(off, -1)
    z, V = np.linalg.eigh(T)
    order = np.argsort(z)
    z, V = z[order], V[:, order]
    return z, be[0] * V[0, :] ** 2


def zwgj(n, a=0.0, b=0.0):
    al, be = _jacobi_recurrence(n, a, b)
    return _golub_welsch(al, be)


def zwglj(n, a=0.0, b=0.0):
    """n points including both end points: modify the last recurrence row so
    that +-1 are nodes (Gautschi's Lobatto construction)."""
    N = n - 2
    al, be = _jacobi_recurrence(n, a, b)
    apb1 = a + b + 1.0
    al[n - 1] = (a - b) / (2.0 * N + apb1 + 1.0)
    be[n - 1] = (4.0 * (N + a + 1.0) * (N + b + 1.0) * (N + apb1)
                 / ((2.0 * N + apb1) * (2.0 * N + apb1 + 1.0) ** 2))
    return _golub_welsch(al, be)


def _radau(n, a, b, end):
    """n-point Gauss-Radau rule with the fixed node ``end`` (+-1): the last
    diagonal entry is chosen so that ``end`` is an eigenvalue."""
    al, be = _jacobi_recurrence(n, a, b)
    # p_{n-1}(end) / p_{n-2}(end) from the monic three-term recurrence
    p0, p1 = 0.0, 1.0
    for k in range(n - 1):
        p0, p1 = p1, (end - al[k]) * p1 - (be[k] * p0 if k > 0 else 0.0)
    al[n - 1] = end - be[n - 1] * p0 / p1
    return _golub_welsch(al, be)


def zwgrjm(n, a=0.0, b=0.0):
    return _radau(n, a, b, -1.0)


def zwgrjp(n, a=0.0, b=0.0):
    return _radau(n, a, b, 1.0)
