"""Pins the oracle's population functions (CPU only).

The division check (population.jl:157-205; cle.jl:240-272) is pinned on a
given state, one check at a time:
  * binomial partitioning conserves molecules: mother + daughter equals the
    count rounded half to even (capture.jl:38 rounding), per gene and species;
  * the daughter's share is Binomial(n, 1/2) (mean, variance, chi-square);
  * halving gives both cells exactly x/2;
  * volumes halve, and each daughter's slot takes the daughter's volume;
  * replacement slots are uniform over 1:N (population.jl:196) or over the
    other N-1 cells (cle.jl:265-268), and a slot chosen twice keeps the later
    divider's daughter (the reference's sequential phase-2 loop).
The volume-scaled transcription and the partitioning together are pinned by
the reference's population <-> dilution equivalence
(test_analytical_moments.jl:448-530): the mean concentrations m/V and p/V of
a Moran population match the dilution model's closed forms.
"""
import math

import numpy as np
import pytest
from scipy import stats

import scrna_amd as sc
from oracle import oracle as orc


def _state(G, N, rng, lo=0.0, hi=60.0):
    m = rng.uniform(lo, hi, size=(G, N))
    p = rng.uniform(lo, hi, size=(G, N))
    return m, p


def test_binomial_partition_conserves_rounded_counts():
    rng = np.random.default_rng(3)
    G, N = 3, 4000
    m, p = _state(G, N, rng, -2.0, 80.0)
    # values on the rounding boundaries: half to even (2.5 -> 2, 3.5 -> 4), and negatives -> 0
    m[0, :8] = [0.5, 1.5, 2.5, 3.5, -0.5, -3.0, 0.0, 1e-9]
    V = rng.uniform(1.0, 3.0, size=N)
    pop = sc.PopulationConfig(cell_num=N, V_div=2.0, partition="binomial", replace="uniform")
    r = orc.divide(pop, m, p, V, seed=11, step=9)
    div = r["dividers"]
    np.testing.assert_array_equal(div, np.flatnonzero(V > 2.0))
    for pre, mom, dau in ((m, r["mother_m"], r["daughter_m"]), (p, r["mother_p"], r["daughter_p"])):
        n = np.maximum(np.array([[np.round(x) for x in row] for row in pre[:, div].T]), 0)   # numpy rounds half to even
        np.testing.assert_array_equal(mom + dau, n)
        assert np.all(dau >= 0) and np.all(mom >= 0)
        assert np.all(dau == np.floor(dau)) and np.all(mom == np.floor(mom))
    # the boundary values of gene 0
    k = {c: i for i, c in enumerate(div)}
    want = {0: 0, 1: 2, 2: 2, 3: 4, 4: 0, 5: 0, 6: 0, 7: 0}
    for c, n in want.items():
        if c in k:
            assert r["mother_m"][k[c], 0] + r["daughter_m"][k[c], 0] == n


def test_binomial_partition_daughter_law():
    """Daughters of n = 40 molecules: Binomial(40, 1/2) by mean, variance and histogram."""
    G, N = 2, 20000
    m = np.full((G, N), 40.0)
    p = np.full((G, N), 7.0)
    V = np.full(N, 2.5)
    pop = sc.PopulationConfig(cell_num=N, V_div=2.0, partition="binomial", replace="uniform")
    r = orc.divide(pop, m, p, V, seed=5, step=0)
    assert r["dividers"].size == N
    for d, n in ((r["daughter_m"].ravel(), 40), (r["daughter_p"].ravel(), 7)):
        se = math.sqrt(n / 4 / d.size)
        assert abs(d.mean() - n / 2) < 5 * se
        assert abs(d.var() / (n / 4) - 1) < 0.05
        ks = np.arange(n + 1)
        obs = np.bincount(d.astype(int), minlength=n + 1)
        exp = stats.binom.pmf(ks, n, 0.5) * d.size
        keep = exp > 5
        chi = ((obs[keep] - exp[keep]) ** 2 / exp[keep]).sum()
        assert stats.chi2.sf(chi, keep.sum() - 1) > 1e-4
    # genes and species draw independent streams
    assert abs(np.corrcoef(r["daughter_m"][:, 0], r["daughter_m"][:, 1])[0, 1]) < 0.05


def test_halving_partition_is_exact():
    rng = np.random.default_rng(8)
    G, N = 4, 500
    m, p = _state(G, N, rng)
    V = rng.uniform(1.0, 3.0, size=N)
    pop = sc.PopulationConfig(cell_num=N, V_div=2.0, partition="halve", replace="exclude_mother")
    r = orc.divide(pop, m, p, V, seed=2, step=4)
    div = r["dividers"]
    np.testing.assert_array_equal(r["mother_m"], m[:, div].T / 2)
    np.testing.assert_array_equal(r["daughter_m"], m[:, div].T / 2)
    np.testing.assert_array_equal(r["daughter_p"], p[:, div].T / 2)


@pytest.mark.parametrize("replace", ["uniform", "exclude_mother"])
def test_replacement_slots_and_volumes(replace):
    rng = np.random.default_rng(21)
    G, N = 2, 50
    m, p = _state(G, N, rng)
    V = rng.uniform(1.0, 3.0, size=N)
    pop = sc.PopulationConfig(cell_num=N, V_div=2.0, partition="binomial", replace=replace)
    r = orc.divide(pop, m, p, V, seed=77, step=19)
    div, rep = r["dividers"], r["replaced"]
    assert div.size > 10 and np.all((rep >= 0) & (rep < N))
    if replace == "exclude_mother":
        assert np.all(rep != div)
    # expected end state, built from the record with the reference's sequential loop
    em = m.copy()
    ep = p.copy()
    eV = V.copy()
    for k, c in enumerate(div):
        em[:, c] = r["mother_m"][k]
        ep[:, c] = r["mother_p"][k]
        eV[c] = V[c] / 2.0
    for k in range(div.size):                      # later dividers overwrite earlier ones
        em[:, rep[k]] = r["daughter_m"][k]
        ep[:, rep[k]] = r["daughter_p"][k]
        eV[rep[k]] = V[div[k]] / 2.0
    np.testing.assert_array_equal(r["m"], em)
    np.testing.assert_array_equal(r["p"], ep)
    np.testing.assert_array_equal(r["V"], eV)
    assert r["V"].size == N                        # Moran: N fixed


@pytest.mark.parametrize("replace", ["uniform", "exclude_mother"])
def test_replacement_is_uniform(replace):
    """Over many checks the replaced slot is uniform over 1:N (or the other N-1 cells)."""
    G, N = 1, 16
    pop = sc.PopulationConfig(cell_num=N, V_div=2.0, partition="halve", replace=replace)
    counts = np.zeros((N, N))                      # [mother, slot]
    m = np.ones((G, N))
    for step in range(1500):
        r = orc.divide(pop, m, m, np.full(N, 2.5), seed=9, step=step)
        np.add.at(counts, (r["dividers"], r["replaced"]), 1)
    total = counts.sum(axis=0)
    exp = np.full(N, counts.sum() / N)
    assert stats.chisquare(total, exp).pvalue > 1e-4
    if replace == "exclude_mother":
        assert np.all(np.diag(counts) == 0)
    else:
        assert np.diag(counts).sum() > 0


def test_population_oracle_invariants():
    """N fixed, volumes inside [V_div/2, V_div*e^{lambda*dt*interval}] after a check
    (test_analytical_moments.jl:384-392); with binomial partitioning and the
    division check after every step (tauleap-style integer counts at each
    check) the state of every cell that divided is whole."""
    net = sc.GeneNetwork(1, np.array([2.0]), np.zeros((1, 1)))
    kin = sc.KineticParams(k_t=1.0, K_d=50.0, n=2.0, mu_m=0.1, mu_p=0.2)
    pop = sc.PopulationConfig(cell_num=1500, growth_rate=0.03, V_div=2.0, V_init=(1.0, 2.0),
                              div_check_interval=1)
    r = orc.cle_population(net, kin, pop, dt=0.1, T=300.0, cell_num=1500, seed=42)
    V = r["V"]
    assert V.size == 1500
    assert V.min() >= 1.0 * 0.99 and V.max() <= 2.0 * math.exp(0.03 * 0.1) * 1.01
    assert np.all(r["m"] >= 0) and np.all(r["p"] >= 0)
    # The cells that divided at the final check hold whole counts: the mother keeps
    # n - d and the daughter d, both integers.  Their volumes are within one growth
    # step of V_div/2 (halved at the last check).
    just = np.isclose(V, V.min()) | (V <= 1.0 * math.exp(0.03 * 0.1) * (1 + 1e-12))
    assert just.sum() > 0
    np.testing.assert_array_equal(r["m"][:, just], np.round(r["m"][:, just]))
    np.testing.assert_array_equal(r["p"][:, just], np.round(r["p"][:, just]))


def _equivalence_case():
    beta, mu_m, mu_p, k_t, lam = 2.0, 0.1, 0.2, 1.0, 0.03
    net = sc.GeneNetwork(1, np.array([beta]), np.zeros((1, 1)))
    kin_nodil = sc.KineticParams(k_t=k_t, K_d=50.0, n=2.0, mu_m=mu_m, mu_p=mu_p)
    kin_dil = sc.KineticParams(k_t=k_t, K_d=50.0, n=2.0, mu_m=mu_m, mu_p=mu_p, dilution=lam)
    return beta, mu_m, mu_p, k_t, lam, net, kin_nodil, kin_dil


@pytest.mark.parametrize("alg,partition", [("cle", "binomial"), ("cle", "halve"), ("tauleap", "binomial")])
def test_population_dilution_equivalence(alg, partition):
    """test_analytical_moments.jl:448-530 on the oracle: N = 3000, T = 500,
    dt = 0.1, division check every 5 steps.  <m/V> = beta/(mu_m+lambda) and
    <p/V> = k_t beta/((mu_m+lambda)(mu_p+lambda)) within 15%; the dilution model's
    means within 5%; no enrichment <m>/(<m/V><V>) within 15%; the dilution
    model's protein Fano 1 + k_t/(mu_m+mu_p+2 lambda) within 15%."""
    beta, mu_m, mu_p, k_t, lam, net, kin_nodil, kin_dil = _equivalence_case()
    N, T, dt = 3000, 500.0, 0.1
    pop = sc.PopulationConfig(cell_num=N, growth_rate=lam, V_div=2.0, V_init=(0.8, 1.2), div_check_interval=5,
                              partition=partition, replace="uniform")
    rel = lambda a, b: abs(a - b) / max(abs(b), 1e-10)
    if alg == "cle":
        dil = orc.cle(net, kin_dil, dt=dt, T=T, cell_num=N, seed=42, threads=4)
        popr = orc.cle_population(net, kin_nodil, pop, dt=dt, T=T, cell_num=N, seed=42)
    else:
        dil = orc.tauleap(net, kin_dil, dt=dt, T=T, cell_num=N, seed=42, threads=4)
        popr = orc.tauleap_population(net, kin_nodil, pop, dt=dt, T=T, cell_num=N, seed=42)
    V = popr["V"]
    c_m = popr["m"][0] / V
    c_p = popr["p"][0] / V
    c_exact = beta / (mu_m + lam)
    cp_exact = k_t * beta / ((mu_m + lam) * (mu_p + lam))
    assert rel(c_m.mean(), c_exact) < 0.15
    assert rel(c_p.mean(), cp_exact) < 0.15
    assert rel(dil["m"][0].mean(), c_exact) < 0.05
    assert rel(dil["p"][0].mean(), cp_exact) < 0.05
    assert abs(popr["m"][0].mean() / (c_m.mean() * V.mean()) - 1.0) < 0.15
    fano = dil["p"][0].var(ddof=1) / dil["p"][0].mean()
    assert rel(fano, 1.0 + k_t / (mu_m + mu_p + 2 * lam)) < 0.15


def test_equivalence_check_separates_dilution_from_none():
    """The 15% band is narrower than the effect it tests: without growth the
    concentration (here the count, V fixed at 1) sits at beta/mu_m, 30% above
    beta/(mu_m+lambda)."""
    beta, mu_m, mu_p, k_t, lam, net, kin_nodil, _ = _equivalence_case()
    r = orc.cle(net, kin_nodil, dt=0.1, T=500.0, cell_num=3000, seed=42, threads=4)
    assert abs(r["m"][0].mean() / (beta / (mu_m + lam)) - 1) > 0.15


def test_ambiguity_taint_marks_without_changing_the_run():
    net = sc.GeneNetwork(2, np.array([3.0, 2.0]), np.array([[0, 1.5], [-1.0, 0]]))
    kin = sc.KineticParams(k_t=2.0, K_d=10.0, n=2.0, mu_m=0.5, mu_p=0.2)
    pop = sc.PopulationConfig(cell_num=400, growth_rate=0.1, V_div=2.0, V_init=(0.8, 1.9), div_check_interval=5)
    base = orc.cle_population(net, kin, pop, dt=0.05, T=20.0, cell_num=400, seed=4)
    none = orc.cle_population(net, kin, pop, dt=0.05, T=20.0, cell_num=400, seed=4, ambiguity=0.0)
    some = orc.cle_population(net, kin, pop, dt=0.05, T=20.0, cell_num=400, seed=4, ambiguity=1e-3)
    every = orc.cle_population(net, kin, pop, dt=0.05, T=20.0, cell_num=400, seed=4, ambiguity=1.0)
    for r in (none, some, every):
        np.testing.assert_array_equal(r["m"], base["m"])
        np.testing.assert_array_equal(r["V"], base["V"])
    assert not none["taint"].any()
    assert 0 < some["taint"].sum() < every["taint"].sum()
    # with a margin that covers every value, every cell descends from some divider or divided itself
    assert every["taint"].mean() > 0.9


def test_growth_noise_law_tauleap():
    """population.jl:130-134: one step multiplies V by exp((lambda + sigma xi) dt):
    log(V1/V0) has mean lambda dt and standard deviation sigma dt."""
    net = sc.GeneNetwork(1, np.array([2.0]), np.zeros((1, 1)))
    kin = sc.KineticParams(k_t=1.0, K_d=50.0, n=2.0, mu_m=0.1, mu_p=0.2)
    lam, sigma, dt, N = 0.05, 0.4, 0.1, 20000
    pop = sc.PopulationConfig(cell_num=N, growth_rate=lam, V_div=1e9, V_init=(0.8, 1.2), div_check_interval=5,
                              growth_noise=sigma)
    v0 = orc.tauleap_population(net, kin, pop, dt=dt, T=0.0, cell_num=N, seed=3)["V"]
    v1 = orc.tauleap_population(net, kin, pop, dt=dt, T=dt, cell_num=N, seed=3)["V"]
    r = np.log(v1 / v0)
    assert abs(r.mean() - lam * dt) < 5 * sigma * dt / np.sqrt(N)
    assert abs(r.std() / (sigma * dt) - 1) < 0.03
    assert stats.kstest((r - lam * dt) / (sigma * dt), "norm").pvalue > 1e-4


def test_growth_noise_ignored_by_cle_and_volume_law_kept():
    """cle.jl:169 grows volumes deterministically whatever growth_noise is; with the
    tau-leap the noisy growth keeps the Moran volume law within the reference's 15%
    (test_analytical_moments.jl:394-406)."""
    net = sc.GeneNetwork(1, np.array([2.0]), np.zeros((1, 1)))
    kin = sc.KineticParams(k_t=1.0, K_d=50.0, n=2.0, mu_m=0.1, mu_p=0.2)
    base = dict(cell_num=800, growth_rate=0.1, V_div=2.0, V_init=(1.0, 2.0), div_check_interval=1)
    a = orc.cle_population(net, kin, sc.PopulationConfig(**base), dt=0.1, T=30.0, cell_num=800, seed=5)
    b = orc.cle_population(net, kin, sc.PopulationConfig(**base, growth_noise=0.5), dt=0.1, T=30.0, cell_num=800,
                           seed=5)
    np.testing.assert_array_equal(a["V"], b["V"])
    np.testing.assert_array_equal(a["m"], b["m"])
    pop = sc.PopulationConfig(cell_num=3000, growth_rate=0.03, V_div=2.0, V_init=(1.0, 2.0), div_check_interval=1,
                              growth_noise=0.05)
    V = orc.tauleap_population(net, kin, pop, dt=0.1, T=1000.0, cell_num=3000, seed=42)["V"]
    assert abs(V.mean() / (2.0 * np.log(2.0)) - 1) < 0.15
    assert abs(np.mean(1.0 / V) / 0.75 - 1) < 0.15
