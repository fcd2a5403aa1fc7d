"""GPU binomial tau-leap against the FP64 oracle and the reference's bands.

The GPU computes propensities in FP64 in the oracle's operation order and
draws from the same per-(cell, step, gene) Philox stream with the same exact
samplers, so counts must agree exactly (for n = 2 every Hill value is an exact
integer square; for other n a 1-ulp difference of pow could in principle move
a draw, which these sizes do not hit).  The statistical bands are the
reference's own (test_analytical_moments.jl, test_gpu_parity.jl).
"""
import math

import numpy as np
import pytest

import scrna_amd as sc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

KIN = sc.KineticParams(k_t=5.0, K_d=10.0, n=2.0, mu_m=1.0, mu_p=0.2)


def both(ref):
    return np.concatenate([ref["m"].T, ref["p"].T], axis=1)


@pytest.mark.parametrize("G,cells,telegraph,n", [(12, 300, True, 2.0), (5, 64, False, 2.0), (20, 129, True, 1.7),
                                                 (1, 1, False, 2.0), (33, 70, True, 2.0)])
def test_counts_match_oracle_exactly(ctx, G, cells, telegraph, n):
    rng = np.random.default_rng(G * 7 + cells)
    if G >= 5:
        net = sc.samplers.mixed_network(G, rng, n_tf=(2, 4), targets_per_tf=(2, 5), n_and=(1, 2), n_or=(1, 2),
                                        telegraph=telegraph, basal_range=(1.0, 6.0))
    else:
        net = sc.GeneNetwork(G, np.array([4.0] * G), np.zeros((G, G)))
    kin = sc.KineticParams(k_t=3.0, K_d=10.0 if n == 2.0 else 8.0, n=n, mu_m=0.7, mu_p=0.3)
    dt, T, seed = 0.05, 5.0, 11
    print(f"tau-leap G={G} cells={cells} telegraph={telegraph} n={n}", flush=True)
    res = sc.simulate(net, sc.BinomialTauLeap(dt), kin, cell_num=cells, T=T, readout="both", seed=seed,
                      snapshot_every=25, full=True)
    ref = orc.tauleap(net, kin, dt=dt, T=T, cell_num=cells, seed=seed, snapshot_every=25, threads=8)
    assert np.array_equal(res.y, both(ref))
    assert np.array_equal(res.snapshots[:, 0].astype(np.float64), ref["snaps_m"])
    assert np.array_equal(res.snapshots[:, 1].astype(np.float64), ref["snaps_p"])
    assert np.all(res.y >= 0) and np.all(res.y == np.round(res.y))


def test_dilution_zero_init_and_readouts(ctx):
    net = sc.GeneNetwork(3, np.array([2.0, 3.0, 1.0]), np.array([[0, 2.0, 0], [0, 0, -1.0], [1.5, 0, 0]]))
    kin = sc.KineticParams(k_t=2.0, K_d=10.0, n=2.0, mu_m=0.4, mu_p=0.1, dilution=0.05)
    ref = orc.tauleap(net, kin, dt=0.1, T=8.0, cell_num=200, seed=4, init="zeros")
    for ro, want in (("mrna", ref["m"].T), ("protein", ref["p"].T), ("both", both(ref))):
        y = sc.simulate(net, sc.BinomialTauLeap(0.1), kin, cell_num=200, T=8.0, readout=ro, seed=4, init="zeros")
        assert np.array_equal(y, want)


def test_segmented_schedule_and_cell_subset(ctx):
    # G = 20 runs 7 waves per workgroup, one workgroup per CU: 256 tiles fit at
    # once, so 20,000 cells (313 tiles) take the segmented schedule.
    rng = np.random.default_rng(2020)
    net = sc.samplers.mixed_network(20, rng, n_tf=(4, 5), targets_per_tf=(3, 6), n_and=(1, 2), n_or=(1, 2),
                                    telegraph=True)
    dt, T, seed = 0.05, 10.0, 21
    big = sc.simulate(net, sc.BinomialTauLeap(dt), KIN, cell_num=20000, T=T, readout="both", seed=seed)
    small = sc.simulate(net, sc.BinomialTauLeap(dt), KIN, cell_num=8192, T=T, readout="both", seed=seed)
    assert np.array_equal(big[:8192], small)
    ref = orc.tauleap(net, KIN, dt=dt, T=T, cell_num=20000, cell_lo=19936, cell_hi=20000, seed=seed, threads=8)
    assert np.array_equal(big[19936:], both(ref))


def test_batch_matches_single_and_capture(ctx):
    rng = np.random.default_rng(9)
    nets = [sc.samplers.mixed_network(G, rng, n_tf=(2, 3), targets_per_tf=(2, 4), n_and=(0, 1), n_or=(0, 1))
            for G in (6, 11, 17)]
    cells = [100, 257, 64]
    out = sc.simulate_gpu_batch(nets, sc.BinomialTauLeap(0.05), KIN, cell_num=cells, T=[3.0, 4.0, 2.5],
                                readout="protein", seed=5, full=True,
                                capture=sc.CaptureModel(efficiency=0.3, efficiency_std=0.2, readout="protein"))
    for i, net in enumerate(nets):
        ref = orc.tauleap(net, KIN, dt=0.05, T=[3.0, 4.0, 2.5][i], cell_num=cells[i], seed=5, net_id=i)
        assert np.array_equal(out[i].y, ref["p"].T)
        want, _ = orc.capture(ref["p"], 0.3, 0.2, 5, net_id=i)
        assert np.array_equal(out[i].counts, want.T)


def test_reference_band_single_gene(ctx):
    """test_analytical_moments.jl:85-93: BinomialTauLeap(0.1), beta=2, mu_m=0.1,
    mu_p=0.2, k_t=1, 10,000 cells, T=500."""
    beta, mu_m, mu_p, k_t = 2.0, 0.1, 0.2, 1.0
    net = sc.GeneNetwork(1, np.array([beta]), np.zeros((1, 1)))
    kin = sc.KineticParams(k_t=k_t, K_d=50.0, n=2.0, mu_m=mu_m, mu_p=mu_p)
    y = sc.simulate(net, sc.BinomialTauLeap(0.1), kin, cell_num=10000, T=500.0, readout="both", seed=42)
    m, p = y[:, 0], y[:, 1]
    rel = lambda a, b: abs(a - b) / abs(b)
    assert rel(m.mean(), beta / mu_m) < 0.03
    assert rel(m.var(ddof=1), beta / mu_m) < 0.15
    assert rel(p.mean(), k_t * beta / (mu_m * mu_p)) < 0.03
    assert rel(p.var(ddof=1) / p.mean(), 1 + k_t / (mu_m + mu_p)) < 0.15


def test_reference_band_dilution(ctx):
    """test_analytical_moments.jl:296-305 (dilution 0.03)."""
    beta, mu_m, mu_p, k_t, dil = 2.0, 0.1, 0.2, 1.0, 0.03
    net = sc.GeneNetwork(1, np.array([beta]), np.zeros((1, 1)))
    kin = sc.KineticParams(k_t=k_t, K_d=50.0, n=2.0, mu_m=mu_m, mu_p=mu_p, dilution=dil)
    y = sc.simulate(net, sc.BinomialTauLeap(0.1), kin, cell_num=10000, T=500.0, readout="both", seed=42)
    m_exact = beta / (mu_m + dil)
    p_exact = k_t * m_exact / (mu_p + dil)
    fano_exact = 1.0 + k_t / (mu_m + mu_p + 2 * dil)
    assert abs(y[:, 0].mean() / m_exact - 1) < 0.03
    assert abs(y[:, 1].mean() / p_exact - 1) < 0.03
    assert abs(y[:, 1].var(ddof=1) / y[:, 1].mean() / fano_exact - 1) < 0.15


@pytest.mark.parametrize("k_t", [0.5, 1.0, 2.0, 4.0])
def test_reference_band_fano_sweep(ctx, k_t):
    """test_analytical_moments.jl:334-339: Fano = 1 + k_t/(mu_m+mu_p) within 15%, dt=0.05."""
    net = sc.GeneNetwork(1, np.array([2.0]), np.zeros((1, 1)))
    kin = sc.KineticParams(k_t=k_t, K_d=50.0, n=2.0, mu_m=0.1, mu_p=0.2)
    y = sc.simulate(net, sc.BinomialTauLeap(0.05), kin, cell_num=15000, T=500.0, readout="protein", seed=77)[:, 0]
    assert abs(y.var(ddof=1) / y.mean() / (1 + k_t / 0.3) - 1) < 0.15


def test_exact_discrete_moments_on_gpu(ctx):
    """The stationary law of the discrete chain: mRNA ~ Poisson(beta dt/(1-e^{-mu_m dt})) exactly."""
    beta, mu_m, mu_p, k_t, dt = 30.0, 3.0, 0.5, 2.0, 0.1
    net = sc.GeneNetwork(1, np.array([beta]), np.zeros((1, 1)))
    kin = sc.KineticParams(k_t=k_t, K_d=10.0, n=2.0, mu_m=mu_m, mu_p=mu_p)
    cells = 200000
    y = sc.simulate(net, sc.BinomialTauLeap(dt), kin, cell_num=cells, T=40.0, readout="both", seed=8)
    a, b = math.exp(-mu_m * dt), math.exp(-mu_p * dt)
    L = beta * dt / (1 - a)
    Q = k_t * dt * L / (1 - b)
    assert abs(y[:, 0].mean() - L) < 4 * math.sqrt(L / cells)
    assert abs(y[:, 0].var(ddof=1) / L - 1) < 4 * math.sqrt(2 / cells)
    assert abs(y[:, 1].mean() / Q - 1) < 0.005


def test_gpu_vs_cpu_band(ctx):
    """test_gpu_parity.jl:33-43: GPU BinomialTauLeap vs CPU mean within 5% (independent streams)."""
    net = sc.GeneNetwork(1, np.array([5.0]), np.zeros((1, 1)))
    kin = sc.KineticParams(k_t=2.0, K_d=50.0, n=2.0, mu_m=0.1, mu_p=0.2)
    y_gpu = sc.simulate(net, sc.BinomialTauLeap(0.1), kin, cell_num=5000, T=200.0, readout="protein", seed=100)
    y_cpu = orc.tauleap(net, kin, dt=0.1, T=200.0, cell_num=5000, seed=4242, threads=8)["p"].T
    assert abs(y_cpu.mean() - y_gpu.mean()) / y_cpu.mean() < 0.05


def test_unsupported_combinations_raise(ctx):
    net = sc.GeneNetwork(2, np.array([1.0, 2.0]), np.zeros((2, 2)))
    with pytest.raises(sc.UnsupportedError):
        sc.simulate(net, sc.BinomialTauLeap(0.1), KIN, cell_num=10, T=1.0, precision="bf16")
    with pytest.raises(sc.UnsupportedError):
        sc.simulate(net, sc.BinomialTauLeap(0.1), KIN, "gpu", sc.PopulationConfig(cell_num=10, partition="halve"),
                    T=1.0)


@pytest.mark.parametrize("telegraph", [False, True])
def test_population_matches_oracle(ctx, telegraph):
    """Tau-leap with Moran population (tauleap_binomial.jl:195-261): volume-scaled
    transcription, intrinsic decay, binomial partitioning, uniform replacement;
    division on the device."""
    rng = np.random.default_rng(17 + telegraph)
    net = sc.samplers.mixed_network(10, rng, n_tf=(2, 3), targets_per_tf=(2, 4), n_and=(1, 1), n_or=(1, 1),
                                    telegraph=telegraph)
    pop = sc.PopulationConfig(cell_num=300, growth_rate=0.05, V_div=2.0, V_init=(0.8, 1.2), div_check_interval=5)
    dt, T, seed = 0.1, 20.0, 13
    res = sc.simulate(net, sc.BinomialTauLeap(dt), KIN, "gpu", pop, T=T, readout="both", seed=seed, full=True)
    ref = orc.tauleap_population(net, KIN, pop, dt=dt, T=T, cell_num=300, seed=seed)
    assert np.array_equal(res.volumes, ref["V"])
    assert np.array_equal(res.y, both(ref))


def test_population_reference_band(ctx):
    """test_analytical_moments.jl:233-252: N cells kept, counts >= 0, mean protein
    above half the no-population value."""
    net = sc.GeneNetwork(1, np.array([2.0]), np.zeros((1, 1)))
    kin = sc.KineticParams(k_t=1.0, K_d=50.0, n=2.0, mu_m=0.1, mu_p=0.2)
    pop = sc.PopulationConfig(cell_num=2000, growth_rate=0.03, V_div=2.0, V_init=(0.8, 1.2), div_check_interval=5)
    y = sc.simulate(net, sc.BinomialTauLeap(0.1), kin, "gpu", pop, T=200.0, readout="protein", seed=42)
    assert y.shape == (2000, 1)
    assert np.all(y >= 0)
    assert y.mean() > 0.5 * 100.0 and np.isfinite(y.mean())


@pytest.mark.parametrize("partition_check", [5, 1])
def test_population_growth_noise_matches_oracle(ctx, partition_check):
    """growth_noise (population.jl:128-139): each cell's volume grows by
    exp((lambda + sigma xi) dt) per step, xi from the cell's growth block.  The
    device's FP64 log / cos / exp may differ from the host's in the last bit, so
    volumes agree to 1e-12; the division decisions and hence every count agree
    exactly."""
    rng = np.random.default_rng(23)
    net = sc.samplers.mixed_network(8, rng, n_tf=(2, 3), targets_per_tf=(2, 4), n_and=(1, 1), n_or=(1, 1),
                                    telegraph=True)
    pop = sc.PopulationConfig(cell_num=400, growth_rate=0.05, V_div=2.0, V_init=(0.8, 1.2),
                              div_check_interval=partition_check, growth_noise=0.3)
    dt, T, seed = 0.1, 30.0, 19
    print(f"tau-leap growth noise, check every {partition_check}", flush=True)
    res = sc.simulate(net, sc.BinomialTauLeap(dt), KIN, "gpu", pop, T=T, readout="both", seed=seed, full=True)
    ref = orc.tauleap_population(net, KIN, pop, dt=dt, T=T, cell_num=400, seed=seed)
    np.testing.assert_allclose(res.volumes, ref["V"], rtol=1e-12)
    assert np.array_equal(res.y, both(ref))
    assert not np.allclose(res.volumes, orc.tauleap_population(
        net, KIN, sc.PopulationConfig(cell_num=400, growth_rate=0.05, V_div=2.0, V_init=(0.8, 1.2),
                                      div_check_interval=partition_check), dt=dt, T=T, cell_num=400,
        seed=seed)["V"])


BIG_NET = sc.GeneNetwork(3, np.array([4.0, 2.0, 3.0]), np.array([[0, 0.5, 0], [0, 0, -0.5], [0.3, 0, 0]]))
BIG_KIN = sc.KineticParams(k_t=2.0e5, K_d=1e7, n=2.0, mu_m=0.02, mu_p=0.2)


def test_counts_above_2_24_stay_exact(ctx):
    """Protein counts near 2e8 (past FP32's 2^24 integer range): the state pool
    holds tau-leap counts as int32 through segment boundaries, division phases,
    capture and the final download, so they agree with the oracle exactly."""
    print("tau-leap counts ~2e8, plain / population / capture", flush=True)
    ref = orc.tauleap(BIG_NET, BIG_KIN, dt=0.05, T=5.0, cell_num=256, seed=3)
    assert ref["p"].max() > 2 ** 24 and np.any(ref["p"] % 16 != 0)
    y = sc.simulate(BIG_NET, sc.BinomialTauLeap(0.05), BIG_KIN, cell_num=256, T=5.0, readout="both", seed=3)
    assert np.array_equal(y, both(ref))

    pop = sc.PopulationConfig(cell_num=200, growth_rate=0.05, V_div=2.0, V_init=(0.8, 1.2), div_check_interval=5)
    res = sc.simulate(BIG_NET, sc.BinomialTauLeap(0.1), BIG_KIN, "gpu", pop, T=20.0, readout="both", seed=5,
                      full=True)
    rp = orc.tauleap_population(BIG_NET, BIG_KIN, pop, dt=0.1, T=20.0, cell_num=200, seed=5)
    assert rp["p"].max() > 2 ** 24
    assert np.array_equal(res.volumes, rp["V"])
    assert np.array_equal(res.y, both(rp))

    cap = sc.CaptureModel(efficiency=0.3, efficiency_std=0.2, readout="protein")
    out = sc.simulate_gpu_batch([BIG_NET], sc.BinomialTauLeap(0.05), BIG_KIN, cell_num=[256], T=[5.0],
                                readout="protein", seed=3, full=True, capture=cap)
    assert np.array_equal(out[0].y, ref["p"].T)
    want, _ = orc.capture(ref["p"], 0.3, 0.2, 3, net_id=0)
    assert np.array_equal(out[0].counts, want.T)
