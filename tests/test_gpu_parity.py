"""GPU CLE parity against the FP64 oracle (run on a gfx950 box).

Tolerances:
  * injected identical noise: max |gpu - oracle| / max(|oracle|, 1) <= 1e-5
    (FP32 state vs FP64 oracle, same update rule cle.jl:78-121);
  * Philox replay (GPU draws its own normals, the oracle regenerates the same
    u32 stream and transforms it in double): <= 2e-4 on the same scale;
  * capture and promoter states: bit-exact (integer work).
"""
import numpy as np
import pytest

import scrna_amd as sc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

KIN = sc.KineticParams(k_t=5.0, K_d=10.0, n=2.0, mu_m=1.0, mu_p=0.2)


def small_net():
    return sc.GeneNetwork(4, np.array([2.0, 1.5, 3.0, 1.0]),
                          np.array([[0, 2.0, 0, 0], [-1.5, 0, 0, 0], [0, 0, 0, 3.0], [0, 0, -2.0, 0]]),
                          [sc.CoopEdge(2, [0, 1], 3.0)], [sc.RedunEdge(3, [0, 1], 2.5)])


def relerr(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0)))


def both(ref):
    return np.concatenate([ref["m"].T, ref["p"].T], axis=1)


@pytest.mark.parametrize("G,cells,telegraph,seed", [(4, 256, False, 1), (20, 300, True, 2), (37, 65, True, 3),
                                                    (1, 1, False, 4), (9, 130, False, 5)])
def test_injected_noise_parity(ctx, G, cells, telegraph, seed):
    rng = np.random.default_rng(seed)
    if G == 4:
        net = small_net()
    else:
        net = sc.samplers.mixed_network(G, rng, n_tf=(min(4, G), min(5, G)), targets_per_tf=(1, 6),
                                        n_and=(1, 2), n_or=(1, 2), telegraph=telegraph)
    dt, T = 0.01, 2.0
    steps = sc.n_steps(T, dt)
    xi = rng.standard_normal((steps, net.G, 2, cells)).astype(np.float32)
    print(f"case G={G} cells={cells} telegraph={telegraph}", flush=True)
    res = sc.simulate(net, sc.CLE(dt), KIN, cell_num=cells, T=T, readout="both", injected_noise=xi,
                      seed=seed, full=True)
    ref = orc.cle(net, KIN, dt=dt, n_steps=steps, cell_num=cells, injected=xi, seed=seed)
    err = relerr(res.y, both(ref))
    print(f"max rel err {err:.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("init", ["steady", "zeros"])
def test_philox_replay_parity(ctx, init):
    rng = np.random.default_rng(11)
    net = sc.samplers.mixed_network(20, rng, n_tf=(4, 5), targets_per_tf=(3, 6), n_and=(1, 2), n_or=(1, 2),
                                    telegraph=True)
    cells, dt, T, seed = 512, 0.01, 5.0, 1234
    res = sc.simulate(net, sc.CLE(dt), KIN, cell_num=cells, T=T, readout="both", seed=seed, init=init, full=True)
    ref = orc.cle(net, KIN, dt=dt, T=T, cell_num=cells, seed=seed, init=init)
    err = relerr(res.y, both(ref))
    print(f"philox replay max rel err {err:.3e}")
    assert err <= 2e-4


def test_cell_subset_replay(ctx):
    """Oracle replays a subset of cells of a large GPU run (counter-based streams)."""
    net = small_net()
    cells, dt, T, seed = 5000, 0.01, 3.0, 99
    y = sc.simulate(net, sc.CLE(dt), KIN, cell_num=cells, T=T, readout="protein", seed=seed)
    ref = orc.cle(net, KIN, dt=dt, T=T, cell_num=cells, cell_lo=4000, cell_hi=4100, seed=seed)
    assert relerr(y[4000:4100], ref["p"].T) <= 2e-4


def test_snapshots_and_moments(ctx):
    rng = np.random.default_rng(5)
    net = sc.samplers.mixed_network(12, rng, n_tf=(3, 4), targets_per_tf=(2, 4), n_and=(1, 1), n_or=(1, 1),
                                    telegraph=True)
    cells, dt, T = 777, 0.01, 1.0
    res = sc.simulate(net, sc.CLE(dt), KIN, cell_num=cells, T=T, readout="both", seed=3, snapshot_every=25,
                      full=True)
    assert res.snapshots.shape == (4, 2, 12, cells)
    np.testing.assert_array_equal(res.snapshots[-1, 0].T.astype(np.float64), res.y[:, :12])
    np.testing.assert_array_equal(res.snapshots[-1, 1].T.astype(np.float64), res.y[:, 12:])
    for k in range(4):
        m = res.snapshots[k, 0].astype(np.float64)
        p = res.snapshots[k, 1].astype(np.float64)
        want = orc.moments(m, p)
        np.testing.assert_allclose(res.moments[k], want, rtol=1e-12, atol=1e-9)
    ref = orc.cle(net, KIN, dt=dt, T=T, cell_num=cells, seed=3, snapshot_every=25)
    assert relerr(res.snapshots[:, 1].astype(np.float64), ref["snaps_p"]) <= 2e-4


def test_capture_bit_exact(ctx):
    rng = np.random.default_rng(8)
    X = np.round(rng.random((1000, 5)) * 100)
    X[:10, 0] = 0.0
    X[10:20, 1] = 2.5          # round half to even
    X[20:30, 2] = 5000.0       # BTRD branch
    cap = sc.CaptureModel(0.1, 0.2, "mrna")
    Y = sc.apply_capture(X, cap, seed=100)
    Yo, _ = orc.capture(X.T, 0.1, 0.2, 100)
    np.testing.assert_array_equal(Y, Yo.T)
    assert np.all(Y <= np.round(X)) and np.all(Y >= 0)


def test_fused_capture_matches_oracle_on_gpu_state(ctx):
    rng = np.random.default_rng(9)
    net = sc.samplers.mixed_network(20, rng, n_tf=(4, 5), targets_per_tf=(3, 6), n_and=(1, 2), n_or=(1, 2),
                                    telegraph=True)
    cap = sc.CaptureModel(0.1, 0.3, "mrna")
    res = sc.simulate(net, sc.CLE(0.01), KIN, cell_num=3000, T=2.0, readout="mrna", capture=cap, seed=77,
                      full=True)
    Yo, _ = orc.capture(res.y.T, 0.1, 0.3, 77)
    np.testing.assert_array_equal(res.counts, Yo.T)


def test_zero_steps_returns_initial_state(ctx):
    net = small_net()
    y = sc.simulate(net, sc.CLE(0.1), KIN, cell_num=70, T=0.0, readout="both")
    m_ss, p_ss = sc.steady_state(net, KIN)
    np.testing.assert_array_equal(y, np.tile(np.concatenate([np.ceil(m_ss), np.ceil(p_ss)]), (70, 1)))


def test_batch_uneven_matches_single(ctx):
    rng = np.random.default_rng(12)
    nets = [sc.samplers.mixed_network(G, rng, n_tf=(2, 3), targets_per_tf=(1, 3), n_and=(0, 1), n_or=(0, 1),
                                      telegraph=True) for G in (5, 23, 64, 3)]
    cells = [100, 257, 64, 1]
    out = sc.simulate_gpu_batch(nets, sc.CLE(0.01), KIN, cell_num=cells, T=1.0, seed=5, full=True)
    for i, net in enumerate(nets):
        ref = orc.cle(net, KIN, dt=0.01, T=1.0, cell_num=cells[i], seed=5, net_id=i)
        assert out[i].y.shape == (cells[i], net.G)
        assert relerr(out[i].y, ref["p"].T) <= 2e-4


def test_invalid_inputs_raise(ctx):
    net = small_net()
    with pytest.raises(ValueError):
        sc.simulate(net, sc.CLE(-0.1), KIN, cell_num=10, T=1.0)
    with pytest.raises(ValueError):
        sc.simulate(net, sc.CLE(0.1), sc.KineticParams(K_d=-1.0), cell_num=10, T=1.0)
