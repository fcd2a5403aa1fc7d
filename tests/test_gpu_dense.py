"""Dense (MFMA) regulation path against the FP64 oracle.

BF16X3 recovers ~16 mantissa bits of the contraction, BF16 keeps 8, FP32
on the matrix cores keeps FP32 products and sums; all are compared with the
oracle on identical noise.  The G=256 check states the
per-gene mean and variance error over a cell subset that the oracle replays
with the identical Philox stream (the workload's acceptance band is 1%).
"""
import numpy as np
import pytest

import scrna_amd as sc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

KIN = sc.KineticParams(k_t=5.0, K_d=10.0, n=2.0, mu_m=1.0, mu_p=0.2)


def relerr(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0)))


@pytest.mark.parametrize("precision,tol", [("bf16x3", 2e-4), ("bf16", 2e-2), ("fp32_mfma", 1e-5)])
@pytest.mark.parametrize("G,cells", [(40, 50), (96, 33), (130, 21), (200, 17)])
def test_dense_injected_parity(ctx, precision, tol, G, cells):
    rng = np.random.default_rng(G)
    net = sc.samplers.mixed_network(G, rng, n_tf=(6, 8), targets_per_tf=(5, 12), n_and=(2, 3), n_or=(2, 3),
                                    a_range=(0.5, 3.0))
    dt, T = 0.005, 1.0
    steps = sc.n_steps(T, dt)
    xi = rng.standard_normal((steps, G, 2, cells)).astype(np.float32)
    print(f"dense {precision} G={G} cells={cells}", flush=True)
    y = sc.simulate(net, sc.CLE(dt), KIN, cell_num=cells, T=T, readout="both", injected_noise=xi, seed=3,
                    precision=precision)
    ref = orc.cle(net, KIN, dt=dt, n_steps=steps, cell_num=cells, injected=xi, seed=3)
    err = relerr(y, np.concatenate([ref["m"].T, ref["p"].T], axis=1))
    print(f"max rel err {err:.3e}")
    assert err <= tol


def test_dense_telegraph_matches_sparse_fp32(ctx):
    """Telegraph promoters in the dense kernel draw the same switch stream as the
    sparse FP32 kernel: promoter states agree exactly, states to the BF16X3 tolerance."""
    rng = np.random.default_rng(3)
    net = sc.samplers.mixed_network(48, rng, n_tf=(4, 5), targets_per_tf=(3, 6), n_and=(1, 2), n_or=(1, 2),
                                    telegraph=True)
    cells, dt, T = 100, 0.01, 2.0
    steps = sc.n_steps(T, dt)
    xi = rng.standard_normal((steps, net.G, 2, cells)).astype(np.float32)
    a = sc.simulate(net, sc.CLE(dt), KIN, cell_num=cells, T=T, readout="both", injected_noise=xi, seed=8,
                    precision="bf16x3")
    ref = orc.cle(net, KIN, dt=dt, n_steps=steps, cell_num=cells, injected=xi, seed=8)
    assert relerr(a, np.concatenate([ref["m"].T, ref["p"].T], axis=1)) <= 2e-4


@pytest.mark.parametrize("precision,mean_tol,var_tol", [("bf16x3", 1e-3, 2e-3), ("bf16", 1e-2, 1e-2),
                                                        ("fp32_valu", 1e-3, 2e-3), ("fp32_mfma", 1e-3, 2e-3)])
def test_g256_precision_moments(ctx, precision, mean_tol, var_tol):
    net = sc.samplers.dense_g256_network()
    cells, dt, steps, seed = 384, 0.005, 300, 17
    T = steps * dt
    res = sc.simulate(net, sc.CLE(dt), KIN, cell_num=cells, T=T, readout="both", seed=seed, precision=precision,
                      full=True)
    ref = orc.cle(net, KIN, dt=dt, n_steps=steps, cell_num=cells, seed=seed, threads=16)
    G = net.G
    for k, want in ((0, ref["m"]), (1, ref["p"])):
        got = res.y[:, k * G:(k + 1) * G].T
        mean_err = np.abs(got.mean(1) - want.mean(1)) / np.maximum(np.abs(want.mean(1)), 1e-3)
        var_err = np.abs(got.var(1) - want.var(1)) / np.maximum(want.var(1), 1e-3)
        print(f"{precision} species {k}: max mean err {mean_err.max():.2e}, max var err {var_err.max():.2e}")
        assert mean_err.max() <= mean_tol
        assert var_err.max() <= var_tol


def test_reference_structured_fp32_matches_oracle(ctx):
    """The reference-structured FP32 baseline (per-step elementwise kernels and
    two rocBLAS SGEMMs, ext:87-108) draws the same Philox stream as the fused
    stepper, so it replays the oracle like the fused path does."""
    rng = np.random.default_rng(31)
    net = sc.samplers.mixed_network(12, rng, n_tf=(3, 4), targets_per_tf=(2, 5), n_and=(1, 2), n_or=(1, 2))
    cells, dt, T, seed = 300, 0.01, 2.0, 41
    y = sc.simulate(net, sc.CLE(dt), KIN, cell_num=cells, T=T, readout="both", seed=seed, precision="fp32_unfused")
    ref = orc.cle(net, KIN, dt=dt, T=T, cell_num=cells, seed=seed)
    err = relerr(y, np.concatenate([ref["m"].T, ref["p"].T], axis=1))
    print(f"unfused max rel err {err:.3e}")
    assert err <= 2e-4


def test_fp32_precision_policy(ctx):
    """Default FP32 picks the matrix-core contraction for the dense G=256 network
    (64 additive in-edges per gene) and the sparse VALU stepper for a sparse
    G=200 network: each is bit-identical to the form it picks."""
    dense = sc.samplers.dense_g256_network()
    kw = dict(cell_num=96, T=0.2, readout="both", seed=5)
    a = sc.simulate(dense, sc.CLE(0.005), KIN, precision="fp32", **kw)
    np.testing.assert_array_equal(a, sc.simulate(dense, sc.CLE(0.005), KIN, precision="fp32_mfma", **kw))
    assert not np.array_equal(a, sc.simulate(dense, sc.CLE(0.005), KIN, precision="fp32_valu", **kw))
    sparse = sc.samplers.mixed_network(200, np.random.default_rng(6), n_tf=(4, 5), targets_per_tf=(3, 6),
                                       n_and=(1, 2), n_or=(1, 2))
    b = sc.simulate(sparse, sc.CLE(0.01), KIN, precision="fp32", **kw)
    np.testing.assert_array_equal(b, sc.simulate(sparse, sc.CLE(0.01), KIN, precision="fp32_valu", **kw))


@pytest.mark.parametrize("precision,tol", [("bf16x3", 2e-4), ("bf16", 2e-2), ("fp32_mfma", 1e-5)])
def test_dense_snapshots_and_moments(ctx, precision, tol):
    """Snapshots of the dense stepper (each lane writes its four genes of one
    cell): against the oracle's snapshots on injected noise, with G = 40 (24
    padding genes in the last block), 37 cells (a ragged tile), the last
    snapshot equal to the readout and the FP64 moments equal to sums of the
    snapshots."""
    rng = np.random.default_rng(11)
    G, cells, dt, T, every = 40, 37, 0.005, 0.6, 30
    net = sc.samplers.mixed_network(G, rng, n_tf=(6, 8), targets_per_tf=(5, 12), n_and=(2, 3), n_or=(2, 3),
                                    a_range=(0.5, 3.0))
    steps = sc.n_steps(T, dt)
    xi = rng.standard_normal((steps, G, 2, cells)).astype(np.float32)
    print(f"dense snapshots {precision}", flush=True)
    res = sc.simulate(net, sc.CLE(dt), KIN, cell_num=cells, T=T, readout="both", injected_noise=xi, seed=3,
                      precision=precision, snapshot_every=every, full=True)
    ref = orc.cle(net, KIN, dt=dt, n_steps=steps, cell_num=cells, injected=xi, seed=3, snapshot_every=every)
    n_snap = steps // every
    assert res.snapshots.shape == (n_snap, 2, G, cells)
    for k in range(n_snap):
        assert relerr(res.snapshots[k, 0].astype(np.float64), ref["snaps_m"][k]) <= tol
        assert relerr(res.snapshots[k, 1].astype(np.float64), ref["snaps_p"][k]) <= tol
    np.testing.assert_array_equal(res.snapshots[-1, 0].T.astype(np.float64), res.y[:, :G])
    np.testing.assert_array_equal(res.snapshots[-1, 1].T.astype(np.float64), res.y[:, G:])
    for k in range(n_snap):
        want = orc.moments(res.snapshots[k, 0].astype(np.float64), res.snapshots[k, 1].astype(np.float64))
        np.testing.assert_allclose(res.moments[k], want, rtol=1e-12, atol=1e-9)



@pytest.mark.parametrize("precision,tol", [("bf16x3", 2e-4), ("fp32_mfma", 1e-5)])
def test_dense_negative_initial_counts(ctx, precision, tol):
    """Negative basal rates start some genes below zero (ceil(beta / mu_m) < 0):
    the dense kernel's clamped variant keeps max(m, 0), max(p, 0) and the
    protein variance's clamp on the first step, as cle.jl:109-120 does."""
    rng = np.random.default_rng(29)
    base = sc.samplers.mixed_network(40, rng, n_tf=(6, 8), targets_per_tf=(5, 12), n_and=(2, 3), n_or=(2, 3),
                                     a_range=(0.5, 3.0))
    basals = np.array(base.basals, dtype=np.float64)
    basals[::5] = -np.abs(basals[::5]) - 0.5
    net = sc.GeneNetwork(base.G, basals, np.array(base.interactions), list(base.cooperative), list(base.redundant))
    cells, dt, T = 40, 0.005, 0.5
    steps = sc.n_steps(T, dt)
    xi = rng.standard_normal((steps, net.G, 2, cells)).astype(np.float32)
    print(f"dense negative init {precision}", flush=True)
    y = sc.simulate(net, sc.CLE(dt), KIN, cell_num=cells, T=T, readout="both", injected_noise=xi, seed=3,
                    precision=precision)
    ref = orc.cle(net, KIN, dt=dt, n_steps=steps, cell_num=cells, injected=xi, seed=3)
    assert np.all(y >= 0)
    assert relerr(y, np.concatenate([ref["m"].T, ref["p"].T], axis=1)) <= tol


@pytest.mark.parametrize("precision,tol", [("fp32_mfma", 2e-4), ("bf16x3", 5e-4), ("bf16", 2e-2)])
def test_dense_telegraph_philox_replay(ctx, precision, tol):
    """Free Philox noise with telegraph promoters (the switching bits come from
    the normal words drawn beside the MFMAs): the oracle replays the same
    stream, so switching decisions agree exactly and states to FP32 rounding."""
    rng = np.random.default_rng(31)
    net = sc.samplers.mixed_network(48, rng, n_tf=(4, 5), targets_per_tf=(3, 6), n_and=(1, 2), n_or=(1, 2),
                                    telegraph=True)
    cells, dt, T = 70, 0.01, 1.5
    print(f"dense telegraph replay {precision}", flush=True)
    y = sc.simulate(net, sc.CLE(dt), KIN, cell_num=cells, T=T, readout="both", seed=12, precision=precision)
    ref = orc.cle(net, KIN, dt=dt, T=T, cell_num=cells, seed=12)
    assert relerr(y, np.concatenate([ref["m"].T, ref["p"].T], axis=1)) <= tol
