"""Pinning the oracle's binomial tau-leap (tauleap_binomial.jl:31-180) without
the reference, which cannot run here (no Julia).

* The Poisson sampler (inversion below 10, PTRS above) against scipy's pmf.
* A constitutive gene is an immigration-death chain with binomial thinning:
  m' = Binomial(m, a) + Poisson(beta dt), a = exp(-mu_m dt).  Its stationary
  law is exactly Poisson(L), L = beta dt / (1 - a), so the mRNA histogram
  pins transcription, decay and their order.  The protein recursion
  p' = Binomial(p, b) + Poisson(k_t dt m') is linear, so its stationary mean,
  variance and covariance with m have closed forms at the actual dt; the
  covariance separates "translation reads the updated m" (tauleap_binomial.jl:153)
  from reading the old one by a factor a.
* Initial counts round the steady state half to even (tauleap_binomial.jl:61-62).
"""
import math

import numpy as np
import pytest
from scipy import stats

import scrna_amd as sc
from oracle import oracle as orc

C3_TAULEAP = 8


def exact_two_stage(beta, k_t, mu_m, mu_p, dt, dil=0.0):
    """Stationary moments of the discrete binomial tau-leap chain of one gene."""
    a = math.exp(-(mu_m + dil) * dt)
    b = math.exp(-(mu_p + dil) * dt)
    kd = k_t * dt
    L = beta * dt / (1.0 - a)                     # mean and variance of m (Poisson)
    Q = kd * L / (1.0 - b)                        # mean of p
    Cmp = kd * L / (1.0 - a * b)                  # Cov(m, p)
    V = (b * (1.0 - b) * Q + kd * L * (1.0 + kd) + 2.0 * a * b * kd * Cmp) / (1.0 - b * b)
    return dict(L=L, Q=Q, C=Cmp, V=V, a=a, b=b)


@pytest.mark.parametrize("lam", [0.05, 0.7, 3.5, 9.99, 10.0, 25.0, 150.0, 2500.0])
def test_poisson_sampler_matches_pmf(lam):
    n = 20000
    x = np.array([orc.poisson(lam, i, 17, C3_TAULEAP, 12345) for i in range(n)])
    assert x.min() >= 0
    se = math.sqrt(lam / n)
    assert abs(x.mean() - lam) < 5 * se
    lo, hi = int(stats.poisson.ppf(1e-4, lam)), int(stats.poisson.ppf(1 - 1e-4, lam))
    edges = np.arange(lo, hi + 2)
    obs, _ = np.histogram(np.clip(x, lo, hi), bins=edges)
    exp = np.diff(stats.poisson.cdf(edges - 1, lam))
    exp[0] += stats.poisson.cdf(lo - 1, lam)
    exp[-1] += stats.poisson.sf(hi, lam)
    exp *= n
    keep = exp >= 5
    chi2 = ((obs[keep] - exp[keep]) ** 2 / exp[keep]).sum()
    dof = max(1, keep.sum() - 1)
    assert stats.chi2.sf(chi2, dof) > 1e-4


def test_poisson_degenerate_draws_nothing():
    assert orc.poisson(0.0, 1, 2, C3_TAULEAP, 9) == 0
    assert orc.poisson(-1.0, 1, 2, C3_TAULEAP, 9) == 0


def test_immigration_death_is_exactly_poisson_and_protein_moments():
    beta, mu_m, mu_p, k_t, dt = 30.0, 3.0, 0.5, 2.0, 0.1
    net = sc.GeneNetwork(1, np.array([beta]), np.zeros((1, 1)))
    kin = sc.KineticParams(k_t=k_t, K_d=10.0, n=2.0, mu_m=mu_m, mu_p=mu_p)
    cells = 20000
    r = orc.tauleap(net, kin, dt=dt, T=40.0, cell_num=cells, seed=77, threads=8)
    m, p = r["m"][0], r["p"][0]
    ex = exact_two_stage(beta, k_t, mu_m, mu_p, dt)
    L = ex["L"]
    # mRNA ~ Poisson(L) exactly: mean, Fano and the histogram
    assert abs(m.mean() - L) < 4 * math.sqrt(L / cells)
    assert abs(m.var(ddof=1) / L - 1.0) < 4 * math.sqrt(2.0 / cells)
    lo, hi = int(stats.poisson.ppf(1e-3, L)), int(stats.poisson.ppf(1 - 1e-3, L))
    edges = np.arange(lo, hi + 2)
    obs, _ = np.histogram(np.clip(m, lo, hi), bins=edges)
    exp = np.diff(stats.poisson.cdf(edges - 1, L))
    exp[0] += stats.poisson.cdf(lo - 1, L)
    exp[-1] += stats.poisson.sf(hi, L)
    exp *= cells
    chi2 = ((obs - exp) ** 2 / exp).sum()
    assert stats.chi2.sf(chi2, len(obs) - 1) > 1e-4
    # protein: mean, variance, covariance at the actual dt
    V, Cmp = ex["V"], ex["C"]
    assert abs(p.mean() - ex["Q"]) < 4 * math.sqrt(V / cells)
    assert abs(p.var(ddof=1) / V - 1.0) < 4 * math.sqrt(2.0 / cells)
    c = np.cov(m, p)[0, 1]
    se_c = math.sqrt((L * V + Cmp ** 2) / cells)
    assert abs(c - Cmp) < 4 * se_c
    # the old-mRNA variant would give a * Cmp, 9 standard errors away here
    assert abs(c - ex["a"] * Cmp) > 4 * se_c


def test_dilution_enters_both_decays():
    """tauleap_binomial.jl:144 and :158 use mu + dilution."""
    beta, mu_m, mu_p, k_t, dt, dil = 10.0, 0.5, 0.3, 1.0, 0.1, 0.2
    net = sc.GeneNetwork(1, np.array([beta]), np.zeros((1, 1)))
    kin = sc.KineticParams(k_t=k_t, K_d=10.0, n=2.0, mu_m=mu_m, mu_p=mu_p, dilution=dil)
    cells = 20000
    r = orc.tauleap(net, kin, dt=dt, T=60.0, cell_num=cells, seed=5, threads=8)
    ex = exact_two_stage(beta, k_t, mu_m, mu_p, dt, dil)
    assert abs(r["m"][0].mean() - ex["L"]) < 4 * math.sqrt(ex["L"] / cells)
    assert abs(r["p"][0].mean() - ex["Q"]) < 4 * math.sqrt(ex["V"] / cells)


def test_initial_counts_round_half_to_even():
    # m_ss = 2.5 -> 2, p_ss = 2.5 * 1.4 / 0.4 = 8.75 -> 9; second gene m_ss = 3.5 -> 4
    net = sc.GeneNetwork(2, np.array([2.5, 3.5]), np.zeros((2, 2)))
    kin = sc.KineticParams(k_t=1.4, K_d=10.0, n=2.0, mu_m=1.0, mu_p=0.4)
    r = orc.tauleap(net, kin, dt=0.1, n_steps=0, cell_num=3, seed=1)
    assert r["m"][:, 0].tolist() == [2.0, 4.0]
    assert r["p"][:, 0].tolist() == [9.0, round(3.5 * 1.4 / 0.4)]
    z = orc.tauleap(net, kin, dt=0.1, n_steps=0, cell_num=2, seed=1, init="zeros")
    assert not z["m"].any() and not z["p"].any()


def test_telegraph_gates_transcription():
    """A gene that starts OFF and never switches ON (k_on = 0) only decays
    (tauleap_binomial.jl:134-137)."""
    net = sc.GeneNetwork(1, np.array([50.0]), np.zeros((1, 1)), k_on=[0.0], k_off=[1.0])
    kin = sc.KineticParams(k_t=1.0, K_d=10.0, n=2.0, mu_m=1.0, mu_p=0.5)
    r = orc.tauleap(net, kin, dt=0.1, T=30.0, cell_num=50, seed=3)
    assert not r["promoter"].any()
    assert r["m"].max() == 0.0
