"""Validation helpers of the reference (src/validate.jl), for comparing the
GPU simulations with other algorithms or with analytical results.

The same four functions and semantics:
  * ks_statistic(x, y)          two-sample KS statistic, validate.jl:14-41;
  * ks_critical_value(n1, n2)   asymptotic critical value, validate.jl:51-61;
  * compare_moments(x, y)       mean / variance relative errors, validate.jl:70-77;
  * convergence_rate(dts, errs) log-log slope, validate.jl:85-95.

ks_statistic keeps the reference's walk exactly: on a tie the x sample
advances first and the difference is taken after every single step, which
for tied (count) data can exceed the textbook statistic that merges ties.
"""
from __future__ import annotations

from typing import NamedTuple

import numpy as np


def ks_statistic(x, y) -> float:
    """Largest |ECDF_x - ECDF_y| over the merged walk (validate.jl:14-41).

    Both ECDFs grow by repeated addition of 1/n in walk order, as in the
    reference, so the value agrees with it to the last bit."""
    sx = np.sort(np.asarray(x, dtype=np.float64), kind="stable")
    sy = np.sort(np.asarray(y, dtype=np.float64), kind="stable")
    nx, ny = sx.size, sy.size
    if nx == 0 or ny == 0:
        raise ValueError("ks_statistic needs two non-empty samples")
    # Walk order: merge with x before y on ties (the reference takes x when sx <= sy).
    vals = np.concatenate([sx, sy])
    src = np.concatenate([np.zeros(nx, np.int8), np.ones(ny, np.int8)])
    order = np.lexsort((src, vals))
    from_x = src[order] == 0
    ecdf_x = np.cumsum(np.where(from_x, 1.0 / nx, 0.0))
    ecdf_y = np.cumsum(np.where(from_x, 0.0, 1.0 / ny))
    return float(np.max(np.abs(ecdf_x - ecdf_y)))


def ks_critical_value(n1: int, n2: int, alpha: float = 0.05) -> float:
    """c(alpha) * sqrt((n1 + n2) / (n1 n2)) with c = 1.63, 1.36, 1.22 for
    alpha <= 0.01, <= 0.05 and above (validate.jl:51-61)."""
    c = 1.63 if alpha <= 0.01 else (1.36 if alpha <= 0.05 else 1.22)
    return c * float(np.sqrt((n1 + n2) / (n1 * n2)))


class MomentComparison(NamedTuple):
    mean_relerr: float
    var_relerr: float
    passed: bool


def compare_moments(x, y, rtol: float = 0.05) -> MomentComparison:
    """Relative errors of mean and sample variance (n - 1), each against the
    larger magnitude of the pair or 1e-10 (validate.jl:70-77).  The reference's
    field `pass` is `passed` here (a Python keyword)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    m1, m2 = float(np.mean(x)), float(np.mean(y))
    v1, v2 = float(np.var(x, ddof=1)), float(np.var(y, ddof=1))
    mean_err = abs(m1 - m2) / max(abs(m1), abs(m2), 1e-10)
    var_err = abs(v1 - v2) / max(abs(v1), abs(v2), 1e-10)
    return MomentComparison(mean_err, var_err, mean_err < rtol and var_err < rtol)


def convergence_rate(dts, errors) -> float:
    """Least-squares slope of log(max(error, 1e-16)) against log(dt)
    (validate.jl:85-95): about 1 for Euler-Maruyama weak error."""
    ld = np.log(np.asarray(dts, dtype=np.float64))
    le = np.log(np.maximum(np.asarray(errors, dtype=np.float64), 1e-16))
    xb, yb = ld.mean(), le.mean()
    return float(np.sum((ld - xb) * (le - yb)) / np.sum((ld - xb) ** 2))
