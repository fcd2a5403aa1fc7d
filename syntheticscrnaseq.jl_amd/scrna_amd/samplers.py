"""Seeded network generators for the benchmark workloads.

The reference's ``sample_network`` (kinetics.jl:105-205) draws from Julia
RNG streams that cannot be reproduced here, and the benchmark workloads
specify their own distributions, so these generators are this project's
own.  Each function documents the workload it builds.  Indices are 0-based.
"""
from __future__ import annotations

import numpy as np

from .kinetics import CaptureModel, KineticParams, PopulationConfig
from .network import CoopEdge, GeneNetwork, RedunEdge


def _pair_edge(rng, G, used, lo, hi):
    """Two distinct sources and a target not already regulated by them."""
    if G < 3:
        return None
    for _ in range(20):
        src = np.sort(rng.choice(G, size=2, replace=False))
        targets = [g for g in range(G) if g not in src]
        t = int(rng.choice(targets))
        if (t, src[0]) in used or (t, src[1]) in used:
            continue
        used.add((t, int(src[0])))
        used.add((t, int(src[1])))
        return t, [int(src[0]), int(src[1])], float(rng.uniform(lo, hi))
    return None


def mixed_network(G: int, rng: np.random.Generator, *, n_tf, targets_per_tf, n_and, n_or,
                  a_range=(1.0, 4.0), gate_range=(2.0, 5.0), basal_range=(0.5, 3.0),
                  telegraph=False, k_on_range=(0.1, 1.0), k_off_range=(0.5, 5.0)) -> GeneNetwork:
    """TF-level sparse network with additive +-U(a_range) edges and AND/OR gates.

    n_tf, targets_per_tf, n_and and n_or are inclusive (lo, hi) integer ranges.
    """
    basals = rng.uniform(*basal_range, size=G)
    A = np.zeros((G, G))
    used = set()
    ntf = int(rng.integers(n_tf[0], n_tf[1] + 1))
    ntf = min(ntf, G)
    tfs = np.sort(rng.choice(G, size=ntf, replace=False))
    for tf in tfs:
        nt = int(rng.integers(targets_per_tf[0], targets_per_tf[1] + 1))
        cand = [g for g in range(G) if g != tf]
        nt = min(nt, len(cand))
        for t in rng.choice(cand, size=nt, replace=False):
            mag = rng.uniform(*a_range)
            A[int(t), int(tf)] = mag if rng.random() < 0.5 else -mag
            used.add((int(t), int(tf)))
    coop, redun = [], []
    for _ in range(int(rng.integers(n_and[0], n_and[1] + 1))):
        e = _pair_edge(rng, G, used, *gate_range)
        if e:
            coop.append(CoopEdge(*e))
    for _ in range(int(rng.integers(n_or[0], n_or[1] + 1))):
        e = _pair_edge(rng, G, used, *gate_range)
        if e:
            redun.append(RedunEdge(*e))
    k_on = k_off = None
    if telegraph:
        k_on = rng.uniform(*k_on_range, size=G)
        k_off = rng.uniform(*k_off_range, size=G)
    return GeneNetwork(G, basals, A, coop, redun, k_on, k_off)


# Kinetics shared by the workloads: K=10, n=2, mu_m=1, mu_p=0.2, k_t=5, no dilution.
WORKLOAD_KINETICS = KineticParams(k_t=5.0, K_d=10.0, n=2.0, mu_m=1.0, mu_p=0.2, dilution=0.0)


def cpu_g5_workload(seed: int = 0):
    """G=5 mixed network, constitutive, 5000 cells, dt=0.01, T=50, zero init,
    capture LogNormal efficiency 0.1 (sigma 0.3)."""
    rng = np.random.default_rng(seed)
    net = mixed_network(5, rng, n_tf=(1, 2), targets_per_tf=(1, 3), n_and=(0, 1), n_or=(0, 1))
    return dict(net=net, kin=WORKLOAD_KINETICS, cells=5000, dt=0.01, T=50.0, init="zeros",
                capture=CaptureModel(0.1, 0.3, "mrna"))


def gpu_g20_workload(seed: int = 1):
    """G=20 mixed network with telegraph promoters on every gene: 100,000 cells,
    5000 steps at dt=0.01, snapshot every 500 steps, capture on the last."""
    rng = np.random.default_rng(seed)
    net = mixed_network(20, rng, n_tf=(4, 5), targets_per_tf=(3, 6), n_and=(1, 2), n_or=(1, 2),
                        telegraph=True)
    return dict(net=net, kin=WORKLOAD_KINETICS, cells=100_000, dt=0.01, T=50.0, snapshot_every=500,
                init="steady", capture=CaptureModel(0.1, 0.3, "mrna"))


def dense_g256_network(seed: int = 2) -> GeneNetwork:
    """G=256 additive network, 25% nonzeros at seeded positions, +-U(0.5,3),
    plus 8 AND and 8 OR edges, constitutive."""
    return dense_network(256, seed)


def dense_network(G: int, seed: int = 2) -> GeneNetwork:
    """Additive network of G genes, 25% nonzeros at seeded positions, +-U(0.5,3),
    plus 8 AND and 8 OR edges, constitutive (the G=256 precision workload's
    construction at any size)."""
    rng = np.random.default_rng(seed)
    basals = rng.uniform(0.5, 3.0, size=G)
    mask = rng.random((G, G)) < 0.25
    np.fill_diagonal(mask, False)
    vals = rng.uniform(0.5, 3.0, size=(G, G)) * np.where(rng.random((G, G)) < 0.5, 1.0, -1.0)
    A = np.where(mask, vals, 0.0)
    used = {(int(i), int(j)) for i, j in zip(*np.nonzero(mask))}
    coop, redun = [], []
    while len(coop) < 8:
        e = _pair_edge(rng, G, used, 2.0, 5.0)
        if e:
            coop.append(CoopEdge(*e))
    while len(redun) < 8:
        e = _pair_edge(rng, G, used, 2.0, 5.0)
        if e:
            redun.append(RedunEdge(*e))
    return GeneNetwork(G, basals, A, coop, redun)


def batch_networks(n: int = 4096, seed: int = 3):
    """n networks, G uniform on 5..64, cells log-uniform on 256..8192, mixed
    regulation and telegraph promoters (population dynamics in the workload)."""
    rng = np.random.default_rng(seed)
    nets, cells = [], []
    for _ in range(n):
        G = int(rng.integers(5, 65))
        if G <= 5:
            tf, tg, ng = (1, 2), (1, 3), (0, 1)
        elif G <= 10:
            tf, tg, ng = (2, 3), (2, 4), (0, 1)
        else:
            tf, tg, ng = (4, 5), (3, 6), (1, 2)
        nets.append(mixed_network(G, rng, n_tf=tf, targets_per_tf=tg, n_and=ng, n_or=ng, telegraph=True))
        cells.append(int(np.exp(rng.uniform(np.log(256), np.log(8192)))))
    pop = PopulationConfig(growth_rate=0.05, V_div=2.0, V_init=(0.8, 1.2), div_check_interval=10,
                           partition="binomial", replace="uniform")
    return nets, cells, pop
