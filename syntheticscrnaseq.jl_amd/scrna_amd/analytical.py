"""Closed-form moments used to check simulations (host-side helpers).

Telegraph moments follow src/analytical.jl:102-119 (Peccoud-Ycart); the
two-stage model moments follow test/test_analytical_moments.jl:30-35 and the
dilution forms kinetics.jl:18-23.  The exact discrete-time Euler-Maruyama
moments are this project's derivation for the unclamped linear CLE.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import gammaln, hyp1f1


def telegraph_mean(a, b, c):
    """E[m] = a c / (a + b) (analytical.jl:102)."""
    return a * c / (a + b)


def telegraph_variance(a, b, c):
    """Var(m) = mean + a b c^2 / ((a+b)^2 (a+b+1)) (analytical.jl:109-112)."""
    return telegraph_mean(a, b, c) + a * b * c ** 2 / ((a + b) ** 2 * (a + b + 1))


def telegraph_fano(a, b, c):
    """F = 1 + b c / ((a+b)(a+b+1)) (analytical.jl:119)."""
    return 1.0 + b * c / ((a + b) * (a + b + 1))


def telegraph_logpmf(n: int, a, b, c):
    """log P(n) in the Kummer form (analytical.jl:35-53)."""
    lp = (-c + n * math.log(c) - gammaln(n + 1) + gammaln(a + n) - gammaln(a)
          - gammaln(a + b + n) + gammaln(a + b))
    return lp + math.log(hyp1f1(b, a + b + n, c))


def telegraph_distribution(a, b, c, nmax=None, tol=1e-12):
    """P(0..nmax), truncated where P < tol * max P (analytical.jl:71-95)."""
    if nmax is None:
        mu = telegraph_mean(a, b, c)
        sd = math.sqrt(telegraph_variance(a, b, c))
        nmax = max(math.ceil(mu + 10 * sd), 100)
    lp = np.array([telegraph_logpmf(n, a, b, c) for n in range(nmax + 1)])
    cut = math.log(tol) + lp.max()
    last = np.nonzero(lp > cut)[0]
    if last.size and last[-1] < lp.size - 1:
        lp = lp[: last[-1] + 1]
    p = np.exp(lp - lp.max())
    return p / p.sum()


def two_stage_moments(beta, k_t, mu_m, mu_p, dilution=0.0):
    """Exact two-stage moments with dilution mu (test_analytical_moments.jl:30-35, 280-282)."""
    mm, mp = mu_m + dilution, mu_p + dilution
    m = beta / mm
    p = k_t * m / mp
    return dict(m_mean=m, m_var=m, p_mean=p, p_fano=1.0 + k_t / (mm + mp), cov_pm=k_t * m / (mm + mp))


def euler_linear_moments(beta, k_t, mu_m, mu_p, dt):
    """Exact stationary mean and covariance of the unclamped Euler-Maruyama CLE
    for one unregulated constitutive gene:

        m' = m + (beta - mu_m m) dt + sqrt(dt (beta + mu_m m)) xi_m
        p' = p + (k_t m' - mu_p p) dt + sqrt(dt (k_t m' + mu_p p)) xi_p

    Linear in (m, p) with state-dependent noise variance linear in the state,
    so first and second moments close; solves the stationary Lyapunov-type
    equations numerically.  Returns dict(m_mean, p_mean, m_var, p_var, cov).
    """
    a = 1.0 - mu_m * dt
    b = 1.0 - mu_p * dt
    m = beta / mu_m
    p = k_t * m / mu_p
    # Second moments: unknowns X = [Vm, Vp, Cmp] (central).
    # m' = a m + beta dt + s_m xi_m, with E[s_m^2] = dt (beta + mu_m m)
    qm = dt * (beta + mu_m * m)
    # p' = b p + k_t dt m' + s_p xi_p, E[s_p^2] = dt (k_t m + mu_p p) (stationary means)
    qp = dt * (k_t * m + mu_p * p)
    # Vm = a^2 Vm + qm
    Vm = qm / (1.0 - a * a)
    # Cov(m', p') = Cov(m', b p + k_t dt m') = b Cov(m', p) + k_t dt Var(m')
    # Cov(m', p) = a Cmp   =>  C = a b C + k_t dt Vm
    Cmp = k_t * dt * Vm / (1.0 - a * b)
    # Var(p') = b^2 Vp + (k_t dt)^2 Vm + 2 b k_t dt Cov(p, m') + qp ; Cov(p, m') = a Cmp
    Vp = ((k_t * dt) ** 2 * Vm + 2.0 * b * k_t * dt * a * Cmp + qp) / (1.0 - b * b)
    # Cov(m', p') measured at the same time: C_same = a b Cmp_prev... Cmp is Cov(m, p) at stationarity.
    return dict(m_mean=m, p_mean=p, m_var=Vm, p_var=Vp, cov=Cmp)


def telegraph_euler_mean(beta, k_on, k_off, mu_m, dt):
    """Stationary mean mRNA of the CLE with discrete-time promoter switching:
    P(on) = q_on / (q_on + q_off) with q = 1 - exp(-k dt), mean = beta P(on)/mu_m
    (unclamped)."""
    q_on = 1.0 - math.exp(-k_on * dt)
    q_off = 1.0 - math.exp(-k_off * dt)
    return beta * (q_on / (q_on + q_off)) / mu_m
