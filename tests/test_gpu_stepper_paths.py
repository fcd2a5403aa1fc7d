"""Stepper code paths that the workload networks do not reach, against the
FP64 oracle on identical injected noise (tolerances as in test_gpu_parity:
1e-5 relative for the FP32 stepper, 2e-4 for the BF16x3 dense kernel).

  * switching probabilities that round to 1 (threshold 2^32: every cell of
    that promoter state flips each step);
  * genes with more in-edges than the stepper keeps in registers (the rest
    stream from the network descriptor), for every genes-per-lane tile;
  * AND / OR gates with three sources and targets with two gates (the
    dense kernel keeps only one two-source gate per gene in registers);
  * negative initial counts (negative basal rates): the first step keeps the
    update's max(m, 0), max(p, 0) clamps, which launches without such a
    network leave out.
"""
import numpy as np
import pytest

import scrna_amd as sc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

KIN = sc.KineticParams(k_t=5.0, K_d=10.0, n=2.0, mu_m=1.0, mu_p=0.2)


def relerr(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0)))


def run_both(net, cells, dt, T, seed, precision="fp32", ctx=None):
    steps = sc.n_steps(T, dt)
    rng = np.random.default_rng(seed)
    xi = rng.standard_normal((steps, net.G, 2, cells)).astype(np.float32)
    y = sc.simulate(net, sc.CLE(dt), KIN, cell_num=cells, T=T, readout="both", injected_noise=xi, seed=seed,
                    precision=precision, ctx=ctx)
    ref = orc.cle(net, KIN, dt=dt, n_steps=steps, cell_num=cells, injected=xi, seed=seed)
    return relerr(y, np.concatenate([ref["m"].T, ref["p"].T], axis=1))


def test_saturated_switching_thresholds(ctx):
    # k*dt = 100: 1 - exp(-100) rounds the u32 threshold up to 2^32.
    A = np.array([[0, 0, 2.0], [0, 0, 0], [-1.5, 0, 0]])
    net = sc.GeneNetwork(3, np.array([2.0, 1.0, 1.5]), A, [], [],
                         k_on=[1e4, 1.5, np.inf], k_off=[2.0, 1e4, np.inf])
    print("saturated thresholds G=3 cells=130", flush=True)
    err = run_both(net, 130, 0.01, 1.0, 5, ctx=ctx)
    print(f"max rel err {err:.3e}")
    assert err <= 1e-5


def wide_network(G, rng, per_gene=(3, 6), three_source=True):
    """Every gene has several additive in-edges; gates with two and three
    sources, and one target carrying both an AND and an OR gate."""
    A = np.zeros((G, G))
    for i in range(G):
        k = int(rng.integers(per_gene[0], per_gene[1] + 1))
        src = rng.choice([j for j in range(G) if j != i], size=k, replace=False)
        A[i, src] = rng.uniform(0.5, 3.0, size=k) * np.where(rng.random(k) < 0.5, 1.0, -1.0)
    t = int(rng.integers(0, G))
    others = [j for j in range(G) if j != t]
    s3 = [int(x) for x in rng.choice(others, size=3, replace=False)]
    s2 = [int(x) for x in rng.choice(others, size=2, replace=False)]
    coop = [sc.CoopEdge(t, sorted(s3) if three_source else sorted(s3[:2]), 3.0)]
    redun = [sc.RedunEdge(t, sorted(s2), 2.0), sc.RedunEdge((t + 1) % G, sorted(s3[:2]), 2.5)]
    return sc.GeneNetwork(G, rng.uniform(0.5, 3.0, size=G), A, coop, redun,
                          k_on=rng.uniform(0.5, 2.0, size=G), k_off=rng.uniform(0.5, 2.0, size=G))


@pytest.mark.parametrize("ch", [2, 4, 8, 16])
def test_overflow_edges_and_wide_gates_every_tile(ch):
    rng = np.random.default_rng(40 + ch)
    net = wide_network(23, rng)
    ctx2 = sc.Context(0)
    try:
        ctx2.set_tile(1, min(64, 16 * ch), ch)      # at most 16 waves per workgroup
        print(f"wide network G=23 tile {ch}", flush=True)
        err = run_both(net, 70, 0.01, 1.0, ch, ctx=ctx2)
    finally:
        ctx2.close()
    print(f"max rel err {err:.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("ch", [8, 16])
def test_many_streamed_edges_wide_tiles(ch):
    """Wide tiles stream in-edges eight per round with a one-edge tail: 9 to 27
    in-edges per gene cover one to three rounds and every tail length."""
    rng = np.random.default_rng(60 + ch)
    net = wide_network(40, rng, per_gene=(9, 27))
    ctx2 = sc.Context(0)
    try:
        ctx2.set_tile(1, 64, ch)
        print(f"many edges G=40 tile {ch}", flush=True)
        err = run_both(net, 70, 0.01, 1.0, 3 + ch, precision="fp32_valu", ctx=ctx2)
    finally:
        ctx2.close()
    print(f"max rel err {err:.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("three_source", [True, False])
def test_dense_gate_layouts(ctx, three_source):
    rng = np.random.default_rng(7 if three_source else 8)
    net = wide_network(40, rng, per_gene=(8, 14), three_source=three_source)
    print(f"dense bf16x3 G=40 three_source={three_source}", flush=True)
    err = run_both(net, 50, 0.005, 0.5, 9, precision="bf16x3", ctx=ctx)
    print(f"max rel err {err:.3e}")
    assert err <= 2e-4


def test_negative_initial_counts_keep_the_clamps(ctx):
    # ceil(beta / mu_m) < 0 for genes 1 and 3: their first step reads max(m, 0) = 0.
    A = np.array([[0, 1.5, 0, 0, 0], [0, 0, -2.0, 0, 0], [0, 0, 0, 1.0, 0], [0, 0, 0, 0, 2.5],
                  [-1.0, 0, 0, 0, 0]])
    net = sc.GeneNetwork(5, np.array([2.0, -2.5, 1.0, -0.7, 3.0]), A, [], [],
                         k_on=[0.5, np.inf, 1.0, 0.3, np.inf], k_off=[2.0, np.inf, 1.5, 4.0, np.inf])
    assert run_both(net, 130, 0.01, 1.0, 41, ctx=ctx) <= 1e-5
    # in a batch the clamped kernel serves the whole launch group (free Philox streams, net_id = index)
    rng = np.random.default_rng(5)
    other = sc.samplers.mixed_network(5, rng, n_tf=(1, 2), targets_per_tf=(1, 3), n_and=(0, 1), n_or=(0, 1),
                                      telegraph=True)
    nets, cells = [net, other], [130, 70]
    out = sc.simulate_gpu_batch(nets, sc.CLE(0.01), KIN, cell_num=cells, T=1.0, readout="both", seed=9,
                                full=True, ctx=ctx)
    for i, (n_, c) in enumerate(zip(nets, cells)):
        ref = orc.cle(n_, KIN, dt=0.01, T=1.0, cell_num=c, seed=9, net_id=i)
        assert relerr(out[i].y, np.concatenate([ref["m"].T, ref["p"].T], axis=1)) <= 2e-4
