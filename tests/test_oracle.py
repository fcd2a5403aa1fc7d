"""Pins the FP64 CPU oracle (CPU only).

Each check is chosen so that a plausible transcription mistake (a dropped
term, a wrong sign or index, sqrt(dt) misplaced, the protein step using the
old mRNA) fails it:
  * Philox4x32-10 against Random123's published known-answer vectors and
    rocRAND's independent implementation;
  * Box-Muller against its defining identities;
  * matrix-form regulation (cle.jl:79-106) against the loop form
    (network.jl:158-193) and a from-definition Python brute force;
  * the noise-free Euler map against its closed form (matrix power);
  * one injected-noise step against the hand-expanded update;
  * stationary moments against the exact discrete-time Euler moments;
  * telegraph switching against the discrete-time promoter chain;
  * the binomial sampler against scipy.stats.binom.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy import stats

import scrna_amd as sc
from oracle import oracle as orc
from scrna_amd import analytical as an

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_philox_random123_kat():
    vecs = json.load(open(os.path.join(GOLD, "philox_kat_random123.json")))["vectors"]
    for v in vecs:
        ctr = [int(x, 16) for x in v["ctr"]]
        key = [int(x, 16) for x in v["key"]]
        assert orc.philox(ctr, key) == [int(x, 16) for x in v["out"]]


def test_philox_rocrand_kat():
    for v in json.load(open(os.path.join(GOLD, "philox_kat_rocrand.json"))):
        assert orc.philox(v["ctr"], v["key"]) == v["out"]


@pytest.mark.parametrize("a,b", [(0, 0), (0xFFFFFFFF, 0xFFFFFFFF), (12345678, 987654321), (1 << 31, 1 << 30)])
def test_normal_pair_identities(a, b):
    z0, z1 = orc.normal_pair(a, b)
    u1 = ((a >> 9) + 0.5) * 2.0 ** -23
    th = (b >> 9) * 2.0 ** -23
    assert math.isclose(z0 * z0 + z1 * z1, -2.0 * math.log(u1), rel_tol=1e-13)
    if abs(z0) + abs(z1) > 0:
        ang = math.atan2(z1, z0) % (2 * math.pi)
        assert math.isclose(ang, (2 * math.pi * th) % (2 * math.pi), abs_tol=1e-12) or th == 0.0


def test_normal_stream_moments():
    rng = np.random.default_rng(0)
    u = rng.integers(0, 2 ** 32, size=(20000, 2), dtype=np.uint64)
    z = np.array([orc.normal_pair(int(a), int(b)) for a, b in u]).ravel()
    assert abs(z.mean()) < 0.02 and abs(z.var() - 1) < 0.03
    assert stats.kstest(z, "norm").pvalue > 1e-3


def _brute_reg(net, kin, p):
    """From-definition regulation (network.jl:112-143 semantics, separate code)."""
    G = net.G
    out = np.zeros(G)
    Kn = kin.K_d ** kin.n
    h = lambda x: x ** kin.n / (x ** kin.n + Kn)
    for i in range(G):
        for j in range(G):
            a = net.interactions[i, j] * net.basals[i]
            if i == j or a == 0:
                continue
            out[i] += a * h(p[j]) if a > 0 else -a * (1.0 - h(p[j]))
        for e in net.cooperative:
            if e.target == i:
                out[i] += e.strength * net.basals[i] * np.prod([h(p[s]) for s in e.sources])
        for e in net.redundant:
            if e.target == i:
                hs = [h(p[s]) for s in e.sources]
                val = hs[0] + hs[1] - hs[0] * hs[1] if len(hs) == 2 else 1 - np.prod([1 - x for x in hs])
                out[i] += e.strength * net.basals[i] * val
    return out


@pytest.mark.parametrize("seed", range(6))
def test_regulation_forms_agree(seed):
    rng = np.random.default_rng(seed)
    G = int(rng.integers(3, 12))
    net = sc.samplers.mixed_network(G, rng, n_tf=(1, 3), targets_per_tf=(1, 4), n_and=(0, 2), n_or=(0, 2))
    kin = sc.KineticParams(k_t=5.0, K_d=10.0, n=float(rng.choice([2.0, 3.0, 10.0, 1.7])), mu_m=1.0, mu_p=0.2)
    for _ in range(5):
        p = rng.uniform(0, 40, size=G)
        p[rng.integers(0, G)] = 0.0
        a = orc.regulation(net, kin, p, "matrix")
        b = orc.regulation(net, kin, p, "loop")
        c = _brute_reg(net, kin, p)
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(a, c, rtol=1e-10, atol=1e-12)


def test_noise_free_euler_closed_form():
    """With xi = 0 the step is the affine map [m',p'] = M [m,p] + c; compare k steps to M^k."""
    beta, k_t, mu_m, mu_p, dt, steps = 3.0, 2.0, 0.5, 0.2, 0.05, 300
    net = sc.GeneNetwork(1, np.array([beta]), np.zeros((1, 1)))
    kin = sc.KineticParams(k_t=k_t, K_d=10.0, n=2.0, mu_m=mu_m, mu_p=mu_p)
    xi = np.zeros((steps, 1, 2, 3), dtype=np.float32)
    r = orc.cle(net, kin, dt=dt, n_steps=steps, cell_num=3, injected=xi, init="zeros")
    a, b = 1 - mu_m * dt, 1 - mu_p * dt
    M = np.array([[a, 0.0], [k_t * dt * a, b]])          # protein uses the updated mRNA
    c = np.array([beta * dt, k_t * dt * beta * dt])
    x = np.zeros(2)
    Mk = np.linalg.matrix_power(M, steps)
    # x_k = sum_{i<k} M^i c
    S = np.linalg.solve(np.eye(2) - M, (np.eye(2) - Mk) @ c)
    np.testing.assert_allclose(r["m"][0], S[0], rtol=1e-12)
    np.testing.assert_allclose(r["p"][0], S[1], rtol=1e-12)
    del x


def test_injected_single_step_by_hand():
    beta, k_t, mu_m, mu_p, dt = 2.0, 5.0, 1.0, 0.2, 0.01
    net = sc.GeneNetwork(1, np.array([beta]), np.zeros((1, 1)))
    kin = sc.KineticParams(k_t=k_t, K_d=10.0, n=2.0, mu_m=mu_m, mu_p=mu_p, dilution=0.1)
    xi = np.array([[[[1.5], [-0.7]]]], dtype=np.float32)
    r = orc.cle(net, kin, dt=dt, n_steps=1, cell_num=1, injected=xi)
    mm, mp = mu_m + 0.1, mu_p + 0.1
    m0 = math.ceil(beta / mm)
    p0 = math.ceil(k_t * (beta / mm) / mp)
    m1 = max(m0 + (beta - mm * m0) * dt + 1.5 * math.sqrt(dt) * math.sqrt(beta + mm * m0), 0.0)
    p1 = max(p0 + (k_t * m1 - mp * p0) * dt + float(np.float32(-0.7)) * math.sqrt(dt) * math.sqrt(k_t * m1 + mp * p0), 0)
    assert math.isclose(r["m"][0, 0], m1, rel_tol=1e-7)
    assert math.isclose(r["p"][0, 0], p1, rel_tol=1e-7)


def test_stationary_moments_match_exact_euler_moments():
    """High-count unregulated gene (clamp inactive): sample moments vs the exact
    discrete-time Euler-Maruyama stationary moments; sampling error only."""
    beta, k_t, mu_m, mu_p, dt = 60.0, 2.0, 0.5, 0.25, 0.2
    net = sc.GeneNetwork(1, np.array([beta]), np.zeros((1, 1)))
    kin = sc.KineticParams(k_t=k_t, K_d=10.0, n=2.0, mu_m=mu_m, mu_p=mu_p)
    cells = 40000
    r = orc.cle(net, kin, dt=dt, n_steps=400, cell_num=cells, seed=5, threads=8)
    m, p = r["m"][0], r["p"][0]
    ex = an.euler_linear_moments(beta, k_t, mu_m, mu_p, dt)
    assert abs(m.mean() / ex["m_mean"] - 1) < 0.01
    assert abs(p.mean() / ex["p_mean"] - 1) < 0.01
    assert abs(m.var() / ex["m_var"] - 1) < 0.04      # ~sqrt(2/N) = 0.7% sampling sd
    assert abs(p.var() / ex["p_var"] - 1) < 0.04
    assert abs(np.cov(m, p)[0, 1] / ex["cov"] - 1) < 0.06
    # the continuous-time limit differs from the Euler moments by O(dt): make sure the test can tell
    ct = an.two_stage_moments(beta, k_t, mu_m, mu_p)
    assert abs(ex["m_var"] / ct["m_var"] - 1) > 0.04


def test_cle_two_stage_band_reference_setting():
    """CPU CLE vs exact two-stage moments (test_analytical_moments.jl:110-121 bands)."""
    beta, mu_m, mu_p, k_t = 2.0, 0.1, 0.2, 1.0
    net = sc.GeneNetwork(1, np.array([beta]), np.zeros((1, 1)))
    kin = sc.KineticParams(k_t=k_t, K_d=50.0, n=2.0, mu_m=mu_m, mu_p=mu_p)
    r = orc.cle(net, kin, dt=0.1, T=500.0, cell_num=10000, seed=42, threads=8)
    m, p = r["m"][0], r["p"][0]
    ex = an.two_stage_moments(beta, k_t, mu_m, mu_p)
    rel = lambda a, b: abs(a - b) / max(abs(b), 1e-10)
    assert rel(m.mean(), ex["m_mean"]) < 0.05
    assert rel(m.var(ddof=1), ex["m_var"]) < 0.20
    assert rel(p.mean(), ex["p_mean"]) < 0.05
    assert rel(p.var(ddof=1) / p.mean(), ex["p_fano"]) < 0.20


def test_dilution_band():
    """test_analytical_moments.jl:306-314: CLE with dilution 0.03."""
    beta, mu_m, mu_p, k_t, mu = 2.0, 0.1, 0.2, 1.0, 0.03
    net = sc.GeneNetwork(1, np.array([beta]), np.zeros((1, 1)))
    kin = sc.KineticParams(k_t=k_t, K_d=50.0, n=2.0, mu_m=mu_m, mu_p=mu_p, dilution=mu)
    r = orc.cle(net, kin, dt=0.1, T=500.0, cell_num=10000, seed=42, threads=8)
    m, p = r["m"][0], r["p"][0]
    ex = an.two_stage_moments(beta, k_t, mu_m, mu_p, dilution=mu)
    assert abs(m.mean() / ex["m_mean"] - 1) < 0.05
    assert abs(p.mean() / ex["p_mean"] - 1) < 0.05
    assert abs(p.var(ddof=1) / p.mean() / ex["p_fano"] - 1) < 0.20


def test_telegraph_promoter_chain():
    """Promoter ON fraction = q_on/(q_on+q_off), q = 1-exp(-k dt) (tauleap_binomial.jl:110-121),
    and the unclamped mean mRNA = beta P(on)/mu_m."""
    k_on, k_off, beta, mu_m, dt = 0.5, 2.0, 40.0, 0.5, 0.05
    net = sc.GeneNetwork(1, np.array([beta]), np.zeros((1, 1)), k_on=[k_on], k_off=[k_off])
    kin = sc.KineticParams(k_t=1.0, K_d=10.0, n=2.0, mu_m=mu_m, mu_p=0.5)
    r = orc.cle(net, kin, dt=dt, n_steps=600, cell_num=40000, seed=3, threads=8)
    q_on, q_off = 1 - math.exp(-k_on * dt), 1 - math.exp(-k_off * dt)
    pon = q_on / (q_on + q_off)
    frac = r["promoter"].mean()
    assert abs(frac - pon) < 4 * math.sqrt(pon * (1 - pon) / 40000)
    # telegraph_mean (analytical.jl:102) with a=k_on/mu, b=k_off/mu, c=beta/mu; the CLE clamps at zero
    # so the sample mean is slightly above the unclamped value
    want = an.telegraph_euler_mean(beta, k_on, k_off, mu_m, dt)
    assert abs(r["m"].mean() / want - 1) < 0.05
    # and the discrete chain is within O((k_on+k_off) dt) of the continuous-time mean
    assert abs(an.telegraph_mean(k_on / mu_m, k_off / mu_m, beta / mu_m) / want - 1) < (k_on + k_off) * dt


def test_cle_switch_decision_is_the_32bit_comparison():
    """The CLE switch decides flip iff X < threshold for the 32-bit uniform
    X = (a & 511) << 23 | (b & 511) << 14 | e, with e the top 14 bits of the gene's
    word of the switch block (Philox pinned by the KATs above).  Checked at every
    boundary of the 18-bit early decisions and on thresholds whose answer hinges on e."""
    rng = np.random.default_rng(0)
    for seed, cell, step, g, c3 in [(99, 5, 7, 6, 0), ((7 << 32) | 3, 123456, 4999, 1, (3 << 4)), (1, 0, 0, 0, 16)]:
        blk = orc.philox([cell, step, g >> 2, c3 | 1], [seed & 0xFFFFFFFF, seed >> 32])
        e = blk[g & 3] >> 18
        thrs = [0, 1, 16383, 16384, 16385, 123456789, (1 << 32) - 1, 1 << 32,
                (5 << 14) + e, (5 << 14) + e + 1, ((1 << 18) - 1 << 14) + e + 1]
        n_rare = 0
        for thr in thrs:
            T = thr >> 14
            for q in sorted({0, 1, 5, T - 1, T, T + 1, (1 << 18) - 1}):
                if not 0 <= q < (1 << 18):
                    continue
                hi = rng.integers(0, 1 << 23, size=2)
                a = int(hi[0]) << 9 | (q >> 9)
                b = int(hi[1]) << 9 | (q & 511)
                want = ((q << 14) | e) < thr
                n_rare += (q << 14) < thr < ((q + 1) << 14)
                assert orc.cle_flip(a, b, thr, seed, cell, step, g, c3) == want, (seed, thr, q)
        assert n_rare >= 3          # decisions that read e were exercised


def test_telegraph_closed_form_kats():
    """README.md:117-120 and test_telegraph_distribution.jl:39-58."""
    assert an.telegraph_mean(2.0, 5.0, 70.0) == pytest.approx(20.0)
    assert an.telegraph_fano(2.0, 5.0, 70.0) == pytest.approx(7.25)
    # test_telegraph_distribution.jl:41-42 asserts |F - 1| < 0.01 here, but analytical.jl:119
    # gives F = 1 + 500*40/(1000*1001) = 1.01998; the near-Poisson limit holds within 0.02.
    assert an.telegraph_fano(500.0, 500.0, 40.0) == pytest.approx(1.0 + 20000.0 / 1001000.0, rel=1e-14)
    assert abs(an.telegraph_fano(500.0, 500.0, 40.0) - 1.0) < 0.02
    for a, b, c in [(2.0, 5.0, 70.0), (0.2, 0.3, 50.0), (0.5, 5.0, 110.0)]:
        p = an.telegraph_distribution(a, b, c)
        assert abs(p.sum() - 1) < 1e-6
        ns = np.arange(p.size)
        mu = (ns * p).sum()
        assert abs(mu - an.telegraph_mean(a, b, c)) < 0.01
        v = ((ns - mu) ** 2 * p).sum()
        assert abs(v / an.telegraph_variance(a, b, c) - 1) < 1e-3


@pytest.mark.parametrize("n,p", [(5, 0.3), (40, 0.1), (1000, 0.1), (20000, 0.03), (300, 0.5), (77, 0.93),
                                 (5000, 0.7), (1, 0.5), (12, 0.999)])
def test_binomial_sampler_vs_scipy(n, p):
    draws = np.array([orc.binomial(n, p, c, 7, (3 << 4) | 4, 42) for c in range(6000)])
    assert draws.min() >= 0 and draws.max() <= n
    mean, var = n * p, n * p * (1 - p)
    assert abs(draws.mean() - mean) < 5 * math.sqrt(var / draws.size) + 1e-9
    if var > 1:
        lo, hi = stats.binom.ppf([0.001, 0.999], n, p)
        edges = np.unique(np.linspace(lo, hi, 12).round())
        obs = np.histogram(draws, bins=np.concatenate([[-0.5], edges[1:-1] - 0.5, [n + 0.5]]))[0]
        cdf = stats.binom.cdf(np.concatenate([edges[1:-1] - 1, [n]]), n, p)
        exp = np.diff(np.concatenate([[0.0], cdf])) * draws.size
        keep = exp > 5
        chi = ((obs[keep] - exp[keep]) ** 2 / exp[keep]).sum()
        assert stats.chi2.sf(chi, keep.sum() - 1) > 1e-4


def test_capture_reference_bands():
    """test_capture.jl:8-30 on the oracle's capture."""
    rng = np.random.default_rng(42)
    X = np.round(rng.random((1000, 5)) * 100)
    Y, pc = orc.capture(X.T, 0.1, 0.2, 100)
    Y = Y.T
    assert np.all(Y <= X) and np.all(Y >= 0)
    rates = Y / np.maximum(X, 1)
    assert abs(rates[X > 10].mean() - 0.1) < 0.05
    Yp, _ = orc.capture(X.T, 1.0, 0.01, 200)
    assert (Yp.T / np.maximum(X, 1)).mean() > 0.95
    Yz, _ = orc.capture(X.T, 1e-10, 0.01, 300)
    assert Yz.sum() < X.sum() * 0.01
    # E[p_cell] = efficiency (mean-corrected lognormal, capture.jl:30)
    _, pc = orc.capture(np.ones((1, 100000)), 0.1, 0.3, 5)
    assert abs(pc.mean() / 0.1 - 1) < 0.01
    assert abs(np.log(pc).std() - 0.3) < 0.005
