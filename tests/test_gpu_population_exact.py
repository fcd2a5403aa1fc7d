"""Exact parity of the GPU population path (pop_divide / pop_place and the
population stepper) against the oracle, with no statistical slack.

* One division, exactly: a plan whose networks reach V_div only at their
  last division check is run twice, once with V_div out of reach.  The
  second run's FP32 state is the state the first run divides, so the
  oracle's division (population.jl:157-205) applied to it must reproduce the
  first run's molecules and volumes bit for bit: Binomial(round(n), 1/2)
  partitions, halving, and which daughter lands in which slot, including
  slots claimed twice.
* Many divisions: multi-network plans with injected noise against
  orc.cle_population.  Volumes are bit-exact.  Molecules agree within 1e-4
  on every cell except those the oracle marks as ambiguous: a value that
  its binomial partition rounded lay within 1e-5 * max(1, |x|) of a
  half-integer, where the FP32 state may round the other way, or the cell
  holds a daughter of such a cell.  The marked share is reported and
  bounded.
* The batched population workload's shape: 16 uneven networks from the batch sampler
  (G 5-64, telegraph, AND/OR), growth rate 0.05, division check every 10
  steps, binomial partitioning, uniform replacement, autotuned tiles, the
  whole run captured in one hipGraph and replayed.
* The population <-> dilution equivalence of test_analytical_moments.jl:448-530
  on the GPU CLE and tau-leap.
"""
import numpy as np
import pytest

import scrna_amd as sc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

KIN = sc.KineticParams(k_t=2.0, K_d=10.0, n=2.0, mu_m=0.5, mu_p=0.2)
AMB = 1e-5


def relerr_rows(a, b):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0), axis=1)


def uneven_nets(seed=4):
    rng = np.random.default_rng(seed)
    specs = [(3, 300, True), (7, 129, False), (20, 69, True), (1, 64, False)]
    nets = []
    for G, _, tele in specs:
        if G == 1:
            nets.append(sc.GeneNetwork(1, np.array([3.0]), np.zeros((1, 1)), k_on=[0.7], k_off=[1.3]))
        else:
            nets.append(sc.samplers.mixed_network(G, rng, n_tf=(1, min(3, G)), targets_per_tf=(1, 3),
                                                  n_and=(0, 1), n_or=(0, 1), telegraph=tele))
    return nets, [c for _, c, _ in specs]


@pytest.mark.parametrize("partition,replace", [("binomial", "uniform"), ("binomial", "exclude_mother"),
                                               ("halve", "uniform")])
def test_last_division_bit_exact(ctx, partition, replace):
    nets, cells = uneven_nets()
    dt, I, lam = 0.05, 5, 0.2
    steps = [20, 15, 25, 10]                        # every horizon ends on a division check
    pops, pre_pops = [], []
    for n_steps, c in zip(steps, cells):
        # no cell reaches V_div = 2 before the last check, about a third divide at it
        hi = 2.0 * np.exp(-lam * dt * (n_steps - I)) * (1 - 1e-9)
        lo = hi - 0.27 if n_steps > I else 1.5
        pops.append(sc.PopulationConfig(cell_num=c, growth_rate=lam, V_div=2.0, V_init=(lo, hi),
                                        div_check_interval=I, partition=partition, replace=replace))
        pre_pops.append(sc.PopulationConfig(cell_num=c, growth_rate=lam, V_div=1e300, V_init=(lo, hi),
                                            div_check_interval=I, partition=partition, replace=replace))
    Ts = [s * dt for s in steps]
    print(f"last-division case {partition}/{replace} steps={steps} cells={cells}", flush=True)
    post = sc.simulate_gpu_batch(nets, sc.CLE(dt), KIN, population=pops, T=Ts, readout="both", seed=31, full=True)
    pre = sc.simulate_gpu_batch(nets, sc.CLE(dt), KIN, population=pre_pops, T=Ts, readout="both", seed=31,
                                full=True)
    n_div_total = 0
    for i, net in enumerate(nets):
        G = net.G
        m0 = pre[i].y[:, :G].T.astype(np.float64)
        p0 = pre[i].y[:, G:].T.astype(np.float64)
        want = orc.divide(pops[i], m0, p0, pre[i].volumes, seed=31, step=steps[i] - 1, net_id=i)
        n_div_total += want["dividers"].size
        np.testing.assert_array_equal(post[i].volumes, want["V"])
        np.testing.assert_array_equal(post[i].y[:, :G].T, want["m"])
        np.testing.assert_array_equal(post[i].y[:, G:].T, want["p"])
        if partition == "binomial":
            assert np.all(post[i].y[want["dividers"]] == np.round(post[i].y[want["dividers"]]))
    assert n_div_total > 100
    # slots claimed by two dividers occur in these plans (the later divider's daughter wins)
    assert any(np.unique(orc.divide(pops[i], pre[i].y[:, :nets[i].G].T, pre[i].y[:, nets[i].G:].T,
                                    pre[i].volumes, seed=31, step=steps[i] - 1, net_id=i)["replaced"],
                         return_counts=True)[1].max() > 1 for i in range(len(nets)))


def _compare_population(res, ref, cells, what):
    np.testing.assert_array_equal(res.volumes, ref["V"])
    want = np.concatenate([ref["m"].T, ref["p"].T], axis=1)
    per_cell = relerr_rows(res.y, want)
    clean = ~ref["taint"]
    print(f"{what}: {int(ref['taint'].sum())}/{cells} cells marked ambiguous, "
          f"max rel err on the rest {per_cell[clean].max():.2e}", flush=True)
    assert per_cell[clean].max() <= 1e-4
    assert ref["taint"].mean() <= 0.10
    return per_cell


@pytest.mark.parametrize("replace", ["uniform", "exclude_mother"])
def test_binomial_population_many_divisions(ctx, replace):
    """Uneven networks, one horizon not a multiple of the check interval."""
    nets, cells = uneven_nets(seed=8)
    dt, I = 0.05, 5
    Ts = [6.0, 4.85, 5.5, 3.0]
    pops = [sc.PopulationConfig(cell_num=c, growth_rate=0.15, V_div=2.0, V_init=(0.9, 1.9), div_check_interval=I,
                                partition="binomial", replace=replace) for c in cells]
    rng = np.random.default_rng(12)
    xis = [rng.standard_normal((sc.n_steps(T, dt), net.G, 2, c)).astype(np.float32)
           for net, c, T in zip(nets, cells, Ts)]
    print(f"binomial population, {replace}, T={Ts}", flush=True)
    out = sc.simulate_gpu_batch(nets, sc.CLE(dt), KIN, population=pops, T=Ts, readout="both", seed=5,
                                injected_noise=xis, full=True)
    for i, net in enumerate(nets):
        ref = orc.cle_population(net, KIN, pops[i], dt=dt, T=Ts[i], cell_num=cells[i], seed=5, net_id=i,
                                 injected=xis[i], ambiguity=AMB)
        _compare_population(out[i], ref, cells[i], f"net {i} G={net.G}")


def test_batch_workload_population_graph_autotuned():
    """The 4096-network population batch at test size: 16 networks of the batch sampler, growth
    0.05, check every 10 steps, binomial partition, uniform replacement; tiles
    from the phase autotuner; the whole run in one hipGraph, replayed twice
    with eager work between, each replay equal to the eager plan and to the
    oracle."""
    nets, cells_full, _ = sc.samplers.batch_networks(16, seed=3)
    rng = np.random.default_rng(40)
    cells = [int(rng.integers(64, 260)) for _ in nets]
    dt, T = 0.01, 3.0
    steps = sc.n_steps(T, dt)
    pops = [sc.PopulationConfig(cell_num=c, growth_rate=0.05, V_div=2.0, V_init=(1.4, 1.95), div_check_interval=10,
                                partition="binomial", replace="uniform") for c in cells]
    xis = [rng.standard_normal((steps, net.G, 2, c)).astype(np.float32) for net, c in zip(nets, cells)]
    tctx = sc.Context(0)
    n_buckets = tctx.autotune(max_G=64, phase_steps=10)
    assert n_buckets > 0
    print(f"batch population graph case: G={[n.G for n in nets]} cells={cells}", flush=True)
    eager = sc.Plan(nets, sc.CLE(dt), sc.samplers.WORKLOAD_KINETICS, population=pops, T=T, readout="both", seed=9,
                    injected_noise=xis, ctx=tctx)
    eager.launch()
    eager.sync()
    base = [eager.download(i) for i in range(len(nets))]
    eager.close()
    plan = sc.Plan(nets, sc.CLE(dt), sc.samplers.WORKLOAD_KINETICS, population=pops, T=T, readout="both", seed=9,
                   injected_noise=xis, ctx=tctx)
    plan.use_graph()
    side = sc.simulate(nets[0], sc.CLE(dt), sc.samplers.WORKLOAD_KINETICS, cell_num=100, T=0.2, seed=1, ctx=tctx)
    for rep in range(2):
        plan.launch()
        plan.sync()
        for i in range(len(nets)):
            got = plan.download(i)
            np.testing.assert_array_equal(got.y, base[i].y)
            np.testing.assert_array_equal(got.volumes, base[i].volumes)
        again = sc.simulate(nets[0], sc.CLE(dt), sc.samplers.WORKLOAD_KINETICS, cell_num=100, T=0.2, seed=1, ctx=tctx)
        np.testing.assert_array_equal(again, side)
    plan.close()
    marked = 0
    for i, net in enumerate(nets):
        ref = orc.cle_population(net, sc.samplers.WORKLOAD_KINETICS, pops[i], dt=dt, T=T, cell_num=cells[i], seed=9,
                                 net_id=i, injected=xis[i], ambiguity=AMB)
        _compare_population(base[i], ref, cells[i], f"net {i} G={net.G}")
        marked += int(ref["taint"].sum())
    print(f"ambiguous cells over the batch: {marked}/{sum(cells)}")
    tctx.close()


@pytest.mark.parametrize("alg,partition", [("cle", "binomial"), ("cle", "halve"), ("tauleap", "binomial")])
def test_population_dilution_equivalence_gpu(ctx, alg, partition):
    """test_analytical_moments.jl:448-530 on the GPU: <m/V>, <p/V> of the Moran
    population against the dilution closed forms (15%), the GPU dilution model
    within 5%, no enrichment within 15%."""
    beta, mu_m, mu_p, k_t, lam = 2.0, 0.1, 0.2, 1.0, 0.03
    net = sc.GeneNetwork(1, np.array([beta]), np.zeros((1, 1)))
    kin0 = sc.KineticParams(k_t=k_t, K_d=50.0, n=2.0, mu_m=mu_m, mu_p=mu_p)
    kind = sc.KineticParams(k_t=k_t, K_d=50.0, n=2.0, mu_m=mu_m, mu_p=mu_p, dilution=lam)
    pop = sc.PopulationConfig(cell_num=3000, growth_rate=lam, V_div=2.0, V_init=(0.8, 1.2), div_check_interval=5,
                              partition=partition, replace="uniform")
    A = sc.CLE(0.1) if alg == "cle" else sc.BinomialTauLeap(0.1)
    dil = sc.simulate(net, A, kind, cell_num=3000, T=500.0, readout="both", seed=42)
    res = sc.simulate(net, A, kin0, population=pop, T=500.0, readout="both", seed=42, full=True)
    V = res.volumes
    c_m, c_p = res.y[:, 0] / V, res.y[:, 1] / V
    rel = lambda a, b: abs(a - b) / b
    c_exact = beta / (mu_m + lam)
    cp_exact = k_t * beta / ((mu_m + lam) * (mu_p + lam))
    print(f"{alg}/{partition}: <m/V> {c_m.mean():.3f} vs {c_exact:.3f}, <p/V> {c_p.mean():.3f} vs {cp_exact:.3f}")
    assert rel(c_m.mean(), c_exact) < 0.15 and rel(c_p.mean(), cp_exact) < 0.15
    assert rel(dil[:, 0].mean(), c_exact) < 0.05 and rel(dil[:, 1].mean(), cp_exact) < 0.05
    assert abs(res.y[:, 0].mean() / (c_m.mean() * V.mean()) - 1) < 0.15
