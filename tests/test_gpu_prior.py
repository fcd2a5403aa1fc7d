"""SBI batches end to end: networks drawn from the native prior (kinetics.jl:105-205)
feed one batched GPU plan, and each network's cells replay in the FP64 oracle."""
import numpy as np
import pytest

import scrna_amd as sc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

KIN = sc.KineticParams(k_t=5.0, K_d=10.0, n=2.0, mu_m=1.0, mu_p=0.2)


def relerr(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0)))


@pytest.mark.parametrize("alg", ["cle", "tauleap"])
def test_prior_batch_replays(ctx, alg):
    nets = sc.prior.sample_networks(12, 8, seed=2024)
    algo = sc.CLE(0.01) if alg == "cle" else sc.BinomialTauLeap(0.01)
    out = sc.simulate_gpu_batch(nets, algo, KIN, cell_num=96, T=1.5, readout="both", seed=17)
    for i, net in enumerate(nets):
        if alg == "cle":
            ref = orc.cle(net, KIN, dt=0.01, T=1.5, cell_num=96, seed=17, net_id=i)
            assert relerr(out[i], np.concatenate([ref["m"].T, ref["p"].T], axis=1)) <= 2e-4
        else:
            ref = orc.tauleap(net, KIN, dt=0.01, T=1.5, cell_num=96, seed=17, net_id=i)
            assert np.array_equal(out[i], np.concatenate([ref["m"].T, ref["p"].T], axis=1))


def test_theta_network_simulates(ctx):
    G = 4
    theta = np.concatenate([np.log([2.0, 1.0, 3.0, 0.5]), np.random.default_rng(1).uniform(-3, 3, G * (G - 1))])
    net = sc.prior.network_from_theta(theta, G)
    y = sc.simulate(net, sc.CLE(0.01), KIN, cell_num=200, T=2.0, readout="protein", seed=3)
    ref = orc.cle(net, KIN, dt=0.01, T=2.0, cell_num=200, seed=3)
    assert relerr(y, ref["p"].T) <= 2e-4
