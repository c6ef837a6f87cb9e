"""Spectral estimates and polynomial-preconditioner coefficients (host side).

Ritz values of the (Jacobi-)preconditioned operator from the PCG scalars
alpha_j, beta_j: PCG is the Lanczos process in the C^-1 inner product, and
its tridiagonal is (Saad, Iterative Methods for Sparse Linear Systems, 2nd
ed., Sec. 6.7.3)

    T_jj     = 1/alpha_j + beta_{j-1}/alpha_{j-1}   (beta_{-1} = 0)
    T_j,j+1  = sqrt(beta_j) / alpha_j

whose extreme eigenvalues approach lambda_min / lambda_max of D^-1 A from
inside.  The alpha/beta are those the solver records per iteration
(the reference's `a` and `b`, CG.cpp:333-334, 368-369).

Polynomial preconditioner: z = p(M) D^-1 r, M = D^-1 A, p of degree k in the
Chebyshev basis on [0, b]: p(l) = sum_j c_j T_j(2 l / b - 1).  The classical
choice minimises max |1 - l p(l)| on [a, b]:
1 - l p(l) = T_{k+1}(s(l)) / T_{k+1}(s(0)), s(l) = (b + a - 2 l) / (b - a).
"""
from __future__ import annotations

import numpy as np


def lanczos_tridiagonal(alpha, beta) -> np.ndarray:
    """Dense m x m tridiagonal from m alphas and (at least m - 1) betas."""
    alpha = np.asarray(alpha, dtype=np.float64)
    beta = np.asarray(beta, dtype=np.float64)
    m = alpha.size
    if m == 0:
        raise ValueError("need at least one alpha")
    if beta.size < m - 1:
        raise ValueError("need len(beta) >= len(alpha) - 1")
    T = np.zeros((m, m))
    for j in range(m):
        T[j, j] = 1.0 / alpha[j] + (beta[j - 1] / alpha[j - 1] if j > 0 else 0.0)
        if j + 1 < m:
            off = np.sqrt(beta[j]) / alpha[j]
            T[j, j + 1] = T[j + 1, j] = off
    return T


def ritz_values(alpha, beta) -> np.ndarray:
    """Eigenvalues (ascending) of the Lanczos tridiagonal of PCG's alpha/beta."""
    return np.linalg.eigvalsh(lanczos_tridiagonal(alpha, beta))


def cheb_residual_coeffs(a: float, b: float, degree: int) -> np.ndarray:
    """Coefficients c_0..c_k (Chebyshev basis on [0, b]) of the p with
    1 - l p(l) = T_{k+1}(s(l)) / T_{k+1}(s(0)) on [a, b], 0 < a < b.

    p has degree k exactly, so interpolation at k + 1 Chebyshev points of
    [0, b] recovers it (up to rounding)."""
    if not (0 < a < b):
        raise ValueError("need 0 < a < b")
    k = int(degree)
    y = np.cos(np.pi * (np.arange(k + 1) + 0.5) / (k + 1))  # nodes in (-1, 1)
    lam = (y + 1.0) * 0.5 * b
    s0 = (b + a) / (b - a)
    s = (b + a - 2.0 * lam) / (b - a)
    tk = np.polynomial.chebyshev.chebval(s, [0] * (k + 1) + [1])
    t0 = np.cosh((k + 1) * np.arccosh(s0))
    p = (1.0 - tk / t0) / lam
    V = np.polynomial.chebyshev.chebvander(y, k)
    return np.linalg.solve(V, p)


def poly_eval(coeffs, b: float, lam) -> np.ndarray:
    """p(lam) for Chebyshev coefficients on [0, b]."""
    y = 2.0 * np.asarray(lam, dtype=np.float64) / b - 1.0
    return np.polynomial.chebyshev.chebval(y, np.asarray(coeffs, dtype=np.float64))


def gershgorin_bound(ptr, index, value) -> float:
    """max_i sum_j |a_ij| / a_ii: an upper bound of the spectrum of D^-1 A."""
    ptr = np.asarray(ptr)
    index = np.asarray(index)
    value = np.asarray(value, dtype=np.float64)
    n = ptr.size - 1
    rows = np.repeat(np.arange(n), np.diff(ptr))
    diag = np.zeros(n)
    dmask = index == rows
    diag[rows[dmask]] = value[dmask]
    absrow = np.bincount(rows, weights=np.abs(value), minlength=n)
    return float(np.max(absrow / diag))
