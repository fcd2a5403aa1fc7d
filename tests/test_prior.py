"""Network prior sampler (kinetics.jl:105-205) and theta constructor
(network.jl:233-256): libscrna's native code against the Python restatement,
and the restatement against the prior's defining properties.

The reference draws from Julia's MersenneTwister through StatsBase and
Distributions; those streams cannot be reproduced here, so parity with the
reference is distributional (ranges, supports, LogNormal basals, pair-use
exclusion, the mixed-regulation guarantee), and bit-exact between library and
restatement on the shared Philox stream.
"""
import math

import numpy as np
import pytest
from scipy import stats

import scrna_amd as sc
from oracle import prior as orp


def _gates(net):
    g = [(0, e.target, tuple(e.sources), e.strength) for e in net.cooperative]
    g += [(1, e.target, tuple(e.sources), e.strength) for e in net.redundant]
    return g


@pytest.mark.parametrize("G", [1, 2, 3, 5, 8, 12, 20, 64])
@pytest.mark.parametrize("regulation", ["mixed", "additive", "cooperative", "redundant"])
def test_library_matches_restatement(G, regulation):
    for index in range(6):
        net = sc.prior.sample_network(G, seed=99, index=index, regulation=regulation)
        basals, A, gates = orp.sample_network(G, 99, index, regulation)
        assert np.array_equal(net.basals, basals)
        assert np.array_equal(net.interactions, A)
        assert _gates(net) == gates                         # AND gates are drawn before OR gates
        assert np.all(np.isinf(net.k_on)) and np.all(np.isinf(net.k_off))


def test_custom_ranges_and_fixed_basals():
    net = sc.prior.sample_network(10, seed=3, index=7, n_tf_range=(1, 1), targets_per_tf_range=range(2, 3),
                                  n_coop_range=(0, 0), n_redun_range=(2, 2), infer_basals=False)
    b, A, gates = orp.sample_network(10, 3, 7, "mixed", n_tf=(1, 1), targets=(2, 2), coop=(0, 0), redun=(2, 2),
                                     infer_basals=False)
    assert np.array_equal(net.interactions, A) and _gates(net) == gates
    assert np.all(net.basals == math.exp(-1.0))
    assert np.count_nonzero(A.any(axis=0)) == 1            # one TF
    assert np.count_nonzero(A) == 2                         # two targets


def test_restatement_has_the_prior_structure():
    G, n = 20, 400
    logb, strengths, ntf, ntarg, ngates = [], [], [], [], []
    for i in range(n):
        b, A, gates = orp.sample_network(G, 5, i)
        logb.extend(np.log(b))
        assert np.all(np.diag(A) == 0)
        cols = np.nonzero(A.any(axis=0))[0]
        ntf.append(len(cols))
        for c in cols:
            ntarg.append(np.count_nonzero(A[:, c]))
        nz = A[A != 0]
        strengths.extend(nz)
        used = {(t, s) for t, s in zip(*np.nonzero(A))}
        for _, t, srcs, st in gates:
            assert len(srcs) == 2 and srcs[0] < srcs[1] and t not in srcs
            for s in srcs:
                assert (t, s) not in used                    # pair exclusion (kinetics.jl:193-195)
                used.add((t, s))
            strengths.append(st)
        ngates.append(len(gates))
    logb = np.array(logb)
    assert abs(logb.mean() + 1.0) < 4 * 0.5 / math.sqrt(len(logb))          # LogNormal(-1, 0.5)
    assert abs(logb.std() - 0.5) < 0.02
    assert stats.kstest((logb + 1.0) / 0.5, "norm").pvalue > 1e-3
    s = np.array(strengths)
    assert np.all((np.abs(s) >= 3.0) & (np.abs(s) <= 10.0))                  # U([-10,-3] u [3,10])
    assert abs(np.mean(s > 0) - 0.5) < 0.05
    assert stats.kstest((np.abs(s) - 3.0) / 7.0, "uniform").pvalue > 1e-3
    assert set(ntf) <= {4, 5} and len(set(ntf)) == 2                          # n_tf_range 4:5 for G > 10
    assert set(ntarg) <= {3, 4, 5, 6} and len(set(ntarg)) == 4               # targets 3:6
    assert min(ngates) >= 1 and max(ngates) <= 4                             # 1:2 AND + 1:2 OR, pairs permitting


def test_mixed_guarantees_a_gate_for_small_networks():
    # G <= 10 draws 0:1 AND and 0:1 OR; mixed adds an OR when both are 0 (kinetics.jl:169-175)
    for i in range(200):
        _, _, gates = orp.sample_network(6, 11, i)
        assert len(gates) >= 1
    for i in range(50):
        _, _, gates = orp.sample_network(2, 11, i)          # G < 3: no pair edge possible
        assert gates == []


def test_networks_are_pure_functions_of_seed_and_index():
    a = sc.prior.sample_network(15, seed=1, index=4)
    b = sc.prior.sample_networks(15, 3, seed=1, first_index=3)[1]
    assert np.array_equal(a.interactions, b.interactions) and np.array_equal(a.basals, b.basals)
    c = sc.prior.sample_network(15, seed=2, index=4)
    assert not np.array_equal(a.basals, c.basals)


def test_theta_constructor():
    G = 3
    theta = np.arange(1, 7, dtype=float)                     # row-major off-diagonal: a12 a13 a21 a23 a31 a32
    net = sc.prior.network_from_theta(theta, G)
    assert np.array_equal(net.interactions, np.array([[0, 1, 2], [3, 0, 4], [5, 6, 0]], dtype=float))
    assert np.all(net.basals == math.exp(-1.0))
    th2 = np.concatenate([[0.0, math.log(2.0), 1.0], theta])
    net2 = sc.prior.network_from_theta(th2, G)
    assert np.allclose(net2.basals, [1.0, 2.0, math.e])
    b, A = orp.network_from_theta(th2, G)
    assert np.array_equal(net2.basals, b) and np.array_equal(net2.interactions, A)
    with pytest.raises(ValueError):
        sc.prior.network_from_theta(np.zeros(5), G)


def test_invalid_prior_arguments():
    with pytest.raises(ValueError):
        sc.prior.sample_network(0, seed=1)
    with pytest.raises(ValueError):
        sc.prior.sample_network(5, seed=1, regulation="sideways")
