"""1-norm condition estimate of an upper-triangular U on the CPU -- test infrastructure (the
checker of the library's sptrsv_condest and the calibration of the condition-scaled generators).

Hager's method in Higham's form: N. J. Higham, "FORTRAN codes for estimating the one-norm of a
real or complex matrix, with applications to condition estimation", ACM TOMS 14 (1988) 381-396,
Algorithm 4.1 (the iteration of LAPACK's xLACON).  The reference has no condition estimate; the
solves used here are the oracle's restatements of its serial solvers (src/core.cpp:4-32).

Pinned in tests/test_oracle.py against numpy's dense inverse: the estimate is a lower bound of
||U^-1||_1 (each iterate is ||U^-1 x||_1 for ||x||_1 = 1) and equals it on the test factors.
"""
from __future__ import annotations

import numpy as np

from oracle import oracle as O


def norm1(U) -> float:
    """||U||_1, the largest column sum of |u_ij|."""
    return float(np.bincount(U.col, weights=np.abs(U.val), minlength=U.n).max())


def inv_norm1(U) -> float:
    """Estimate of ||U^-1||_1 (a lower bound) from solves with U and U^T."""
    n = U.n
    cp, ri, cv = O.transpose(U.rowptr, U.col, U.val)
    cp = cp.astype(np.uint64)

    def solve(x):  # U^-1 x
        return O.base_gauss_upper(U.rowptr, U.col, U.val, x)[0]

    def solve_t(x):  # U^-T x: U^T is lower triangular, its CSR is U's CSC with the diagonal last
        return O.base_gauss_lower(cp, ri, cv, x)[0]

    def sign(v):
        return np.where(v >= 0, 1.0, -1.0)

    y = solve(np.full(n, 1.0 / n))
    est = float(np.abs(y).sum())
    if n == 1:
        return est
    sg = sign(y)
    z = solve_t(sg)
    j = int(np.argmax(np.abs(z)))
    for _ in range(2, 6):
        x = np.zeros(n)
        x[j] = 1.0
        y = solve(x)
        old, est = est, float(np.abs(y).sum())
        s2 = sign(y)
        if np.array_equal(s2, sg) or est <= old:
            break
        sg = s2
        z = solve_t(sg)
        jlast, j = j, int(np.argmax(np.abs(z)))
        if z[jlast] == abs(z[j]):
            break
    i = np.arange(n)
    x = np.where(i % 2 == 0, 1.0, -1.0) * (1.0 + i / (n - 1))
    return max(est, 2.0 * float(np.abs(solve(x)).sum()) / (3.0 * n))


def kappa1(U) -> float:
    """Estimate of kappa_1(U) = ||U||_1 ||U^-1||_1."""
    return norm1(U) * inv_norm1(U)
