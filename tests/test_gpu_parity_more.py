"""More GPU parity cases: Hill coefficients other than 2, dilution, readout
layouts, and the statistical bands of the reference's own GPU test
(test_gpu_parity.jl:12-30) and moment tests (test_analytical_moments.jl)."""
import numpy as np
import pytest

import scrna_amd as sc
from oracle import oracle as orc
from scrna_amd import analytical as an

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0)))


@pytest.mark.parametrize("n,K", [(10.0, 61.47011276535837), (1.7, 8.0), (3.0, 12.0), (1.0, 5.0)])
def test_hill_coefficients_injected(ctx, n, K):
    """Default KineticParams use n = 10 (kinetics.jl:44); the kernel evaluates x^n as
    exp2(n log2 x) for n != 2."""
    rng = np.random.default_rng(int(n * 10))
    net = sc.samplers.mixed_network(15, rng, n_tf=(3, 4), targets_per_tf=(2, 5), n_and=(1, 2), n_or=(1, 2),
                                    basal_range=(2.0, 8.0))
    kin = sc.KineticParams(k_t=4.0, K_d=K, n=n, mu_m=0.5, mu_p=0.2)
    cells, dt, T = 128, 0.01, 3.0
    steps = sc.n_steps(T, dt)
    xi = rng.standard_normal((steps, net.G, 2, cells)).astype(np.float32)
    y = sc.simulate(net, sc.CLE(dt), kin, cell_num=cells, T=T, readout="both", injected_noise=xi, seed=1)
    ref = orc.cle(net, kin, dt=dt, n_steps=steps, cell_num=cells, injected=xi)
    err = relerr(y, np.concatenate([ref["m"].T, ref["p"].T], axis=1))
    print(f"n={n}: max rel err {err:.3e}")
    assert err <= 2e-5


def test_dilution_and_readouts(ctx):
    net = sc.GeneNetwork(3, np.array([2.0, 3.0, 1.0]), np.array([[0, 2.0, 0], [0, 0, -1.0], [1.5, 0, 0]]))
    kin = sc.KineticParams(k_t=2.0, K_d=10.0, n=2.0, mu_m=0.4, mu_p=0.1, dilution=0.05)
    cells, dt, T = 200, 0.02, 2.0
    steps = sc.n_steps(T, dt)
    xi = np.random.default_rng(2).standard_normal((steps, 3, 2, cells)).astype(np.float32)
    ref = orc.cle(net, kin, dt=dt, n_steps=steps, cell_num=cells, injected=xi)
    for ro, want in (("mrna", ref["m"].T), ("protein", ref["p"].T),
                     ("both", np.concatenate([ref["m"].T, ref["p"].T], axis=1))):
        y = sc.simulate(net, sc.CLE(dt), kin, cell_num=cells, T=T, readout=ro, injected_noise=xi)
        assert y.shape == want.shape
        assert relerr(y, want) <= 1e-5


def test_reference_gpu_vs_cpu_band(ctx):
    """test_gpu_parity.jl:12-30: GPU CLE vs CPU CLE, G=1, beta=5, k_t=2, K_d=50, n=2,
    5000 cells, T=200, dt=0.1: mean within 5%, variance within 15% (independent streams)."""
    net = sc.GeneNetwork(1, np.array([5.0]), np.zeros((1, 1)))
    kin = sc.KineticParams(k_t=2.0, K_d=50.0, n=2.0, mu_m=0.1, mu_p=0.2)
    y_gpu = sc.simulate(net, sc.CLE(0.1), kin, cell_num=5000, T=200.0, readout="protein", seed=42)
    y_cpu = orc.cle(net, kin, dt=0.1, T=200.0, cell_num=5000, seed=4242, threads=8)["p"].T
    assert abs(y_cpu.mean() - y_gpu.mean()) / y_cpu.mean() < 0.05
    assert abs(y_cpu.var(ddof=1) - y_gpu.var(ddof=1)) / y_cpu.var(ddof=1) < 0.15


def test_two_stage_moments_on_gpu(ctx):
    """test_analytical_moments.jl:110-121 bands, plus the exact Euler moments at the same dt."""
    beta, mu_m, mu_p, k_t = 2.0, 0.1, 0.2, 1.0
    net = sc.GeneNetwork(1, np.array([beta]), np.zeros((1, 1)))
    kin = sc.KineticParams(k_t=k_t, K_d=50.0, n=2.0, mu_m=mu_m, mu_p=mu_p)
    y = sc.simulate(net, sc.CLE(0.1), kin, cell_num=20000, T=500.0, readout="both", seed=11)
    m, p = y[:, 0], y[:, 1]
    ex = an.two_stage_moments(beta, k_t, mu_m, mu_p)
    assert abs(m.mean() / ex["m_mean"] - 1) < 0.05
    assert abs(m.var(ddof=1) / ex["m_var"] - 1) < 0.20
    assert abs(p.mean() / ex["p_mean"] - 1) < 0.05
    assert abs(p.var(ddof=1) / p.mean() / ex["p_fano"] - 1) < 0.20


def test_telegraph_moments_on_gpu(ctx):
    """Unregulated telegraph gene: mean against the discrete-time chain and the
    Fano factor against analytical.jl:119 within the dt bias and sampling error."""
    # fast switching and a high mean keep the clamp at zero out of play
    k_on, k_off, beta, mu_m, dt = 5.0, 5.0, 200.0, 1.0, 0.005
    net = sc.GeneNetwork(1, np.array([beta]), np.zeros((1, 1)), k_on=[k_on], k_off=[k_off])
    kin = sc.KineticParams(k_t=1.0, K_d=10.0, n=2.0, mu_m=mu_m, mu_p=1.0)
    y = sc.simulate(net, sc.CLE(dt), kin, cell_num=100000, T=20.0, readout="mrna", seed=5)[:, 0]
    want_mean = an.telegraph_euler_mean(beta, k_on, k_off, mu_m, dt)
    assert abs(y.mean() / want_mean - 1) < 0.02
    fano = y.var() / y.mean()
    want_fano = an.telegraph_fano(k_on / mu_m, k_off / mu_m, beta / mu_m)
    assert abs(fano / want_fano - 1) < 0.08


def test_large_network_reference_structured_path(ctx):
    """G > 256 on the reference-structured FP32 path (precision "fp32_unfused":
    the reference GPU CLE's per-step kernels, rocBLAS products for any G as
    ext:93-94): the oracle replays a subset of its cells, snapshots and the
    fused capture; telegraph promoters are refused on that path with a message."""
    rng = np.random.default_rng(30)
    net = sc.samplers.mixed_network(300, rng, n_tf=(20, 30), targets_per_tf=(4, 12), n_and=(2, 3), n_or=(2, 3))
    kin = sc.samplers.WORKLOAD_KINETICS
    cells, dt, T, seed = 1500, 0.01, 1.0, 77
    print(f"G=300, {cells} cells, {sc.n_steps(T, dt)} steps", flush=True)
    res = sc.simulate(net, sc.CLE(dt), kin, cell_num=cells, T=T, readout="both", seed=seed, snapshot_every=50,
                      capture=sc.CaptureModel(0.1, 0.3, "mrna"), full=True, precision="fp32_unfused")
    ref = orc.cle(net, kin, dt=dt, T=T, cell_num=cells, cell_lo=0, cell_hi=12, seed=seed, threads=4)
    want = np.concatenate([ref["m"].T, ref["p"].T], axis=1)
    err = float(np.max(np.abs(res.y[:12] - want) / np.maximum(np.abs(want), 1.0)))
    print(f"max rel err {err:.2e}")
    assert err <= 5e-5
    assert res.snapshots.shape == (2, 2, 300, cells)
    np.testing.assert_array_equal(res.snapshots[-1, 1].T.astype(np.float64), res.y[:, 300:])
    Yo, _ = orc.capture(res.y[:, :300].T, 0.1, 0.3, seed)
    np.testing.assert_array_equal(res.counts, Yo.T)
    tele = sc.GeneNetwork(300, net.basals, net.interactions, net.cooperative, net.redundant,
                          k_on=np.full(300, 0.5), k_off=np.full(300, 1.0))
    with pytest.raises(sc.UnsupportedError, match="reference-structured"):
        sc.simulate(tele, sc.CLE(dt), kin, cell_num=10, T=0.1, precision="fp32_unfused")


def test_more_networks_than_one_grid_row(ctx):
    """66,000 one-gene networks in one plan: the capture and moment launches
    index networks by blockIdx.y (at most 65,535) and run in chunks.  The
    networks past the first chunk are checked against the oracle's capture of
    their own state and against FP64 moments."""
    n = 66000
    nets = [sc.GeneNetwork(1, np.array([1.0 + (i % 7)]), np.zeros((1, 1))) for i in range(n)]
    cap = sc.CaptureModel(efficiency=0.3, efficiency_std=0.2, readout="mrna")
    print(f"{n} networks, capture and moments", flush=True)
    kin = sc.KineticParams(k_t=5.0, K_d=10.0, n=2.0, mu_m=1.0, mu_p=0.2)
    out = sc.simulate_gpu_batch(nets, sc.CLE(0.01), kin, cell_num=3, T=0.05, readout="mrna", seed=4, capture=cap,
                                full=True)
    for i in (0, 65534, 65535, 65536, n - 1):
        r = out[i]
        want, _ = orc.capture(r.y.T, 0.3, 0.2, 4, net_id=i)
        np.testing.assert_array_equal(r.counts, want.T)
        m = r.y[:, 0].astype(np.float32).astype(np.float64)
        assert r.moments.shape[-1] == 5
        np.testing.assert_allclose(r.moments[0, 0, 0], m.sum(), rtol=1e-12)
