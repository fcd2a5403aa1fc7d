"""The wide path: CLE networks above the fused steppers' 256 genes, one library
GEMM (FP32, BF16 or BF16x3 products) and one fused elementwise kernel per
step.  Against the FP64 oracle on identical injected noise (bars of the dense
forms: 1e-5 FP32, 2e-4 BF16x3, 2e-2 BF16, relative to max(|x|, 1)), on the free
Philox stream for a replayed cell subset, with telegraph promoters, an odd G
(the last gene pair holds one gene), snapshots, capture and the precision
policy ("fp32" picks the FP32 form bit for bit).
"""
import numpy as np
import pytest

import scrna_amd as sc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

KIN = sc.KineticParams(k_t=5.0, K_d=10.0, n=2.0, mu_m=1.0, mu_p=0.2)


def relerr(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0)))


def network(G, seed, telegraph=False):
    rng = np.random.default_rng(seed)
    A = np.where(rng.random((G, G)) < 0.08, rng.uniform(0.5, 3.0, (G, G)), 0.0)
    A *= np.where(rng.random((G, G)) < 0.5, 1.0, -1.0)
    np.fill_diagonal(A, 0.0)
    coop = [sc.CoopEdge(int(t), sorted(int(x) for x in rng.choice([j for j in range(G) if j != t], 2, replace=False)),
                        float(rng.uniform(2.0, 5.0))) for t in rng.choice(G, 3, replace=False)]
    redun = [sc.RedunEdge(int(t), sorted(int(x) for x in rng.choice([j for j in range(G) if j != t], 3, replace=False)),
                          float(rng.uniform(2.0, 5.0))) for t in rng.choice(G, 3, replace=False)]
    k_on = rng.uniform(0.1, 1.0, G) if telegraph else None
    k_off = rng.uniform(0.5, 5.0, G) if telegraph else None
    return sc.GeneNetwork(G, rng.uniform(0.5, 3.0, G), A, coop, redun, k_on=k_on, k_off=k_off)


@pytest.mark.parametrize("G,telegraph,cells", [(300, False, 70), (301, True, 72), (520, False, 64), (263, True, 35)])
def test_wide_path_injected_noise(ctx, G, telegraph, cells):
    """Cell counts that are multiples of 4 take the kernel's 4-cells-per-lane form, the others one cell per lane."""
    net = network(G, G, telegraph)
    dt, T, seed = 0.01, 0.4, 5
    steps = sc.n_steps(T, dt)
    xi = np.random.default_rng(seed).standard_normal((steps, G, 2, cells)).astype(np.float32)
    ref = orc.cle(net, KIN, dt=dt, n_steps=steps, cell_num=cells, injected=xi, seed=seed, threads=4)
    want = np.concatenate([ref["m"].T, ref["p"].T], axis=1)
    for precision, bar in (("fp32", 1e-5), ("bf16x3", 2e-4), ("bf16", 2e-2)):
        print(f"wide G={G} telegraph={telegraph} cells={cells} {precision}", flush=True)
        y = sc.simulate(net, sc.CLE(dt), KIN, cell_num=cells, T=T, readout="both", injected_noise=xi, seed=seed,
                        precision=precision, ctx=ctx)
        err = relerr(y, want)
        print(f"max rel err {err:.3e}")
        assert err <= bar, precision


def test_wide_path_philox_snapshots_capture(ctx):
    G, cells, dt, T, seed = 300, 1500, 0.01, 1.0, 77
    net = network(G, 9, telegraph=True)
    print(f"wide G={G} {cells} cells, {sc.n_steps(T, dt)} steps, free noise", flush=True)
    res = sc.simulate(net, sc.CLE(dt), KIN, cell_num=cells, T=T, readout="both", seed=seed, snapshot_every=50,
                      capture=sc.CaptureModel(0.1, 0.3, "mrna"), full=True, ctx=ctx)
    again = sc.simulate(net, sc.CLE(dt), KIN, cell_num=cells, T=T, readout="both", seed=seed, precision="fp32_mfma",
                        ctx=ctx)
    np.testing.assert_array_equal(res.y, again)         # the policy's choice, and repeatable
    ref = orc.cle(net, KIN, dt=dt, T=T, cell_num=cells, cell_lo=0, cell_hi=12, seed=seed, threads=4)
    want = np.concatenate([ref["m"].T, ref["p"].T], axis=1)
    err = relerr(res.y[:12], want)
    print(f"max rel err {err:.2e}")
    assert err <= 2e-4
    assert res.snapshots.shape == (2, 2, G, cells)
    np.testing.assert_array_equal(res.snapshots[-1, 1].T.astype(np.float64), res.y[:, G:])
    Yo, _ = orc.capture(res.y[:, :G].T, 0.1, 0.3, seed)
    np.testing.assert_array_equal(res.counts, Yo.T)


def test_wide_path_refusals(ctx):
    net = network(300, 3)
    pop = sc.PopulationConfig(cell_num=64, growth_rate=0.05, V_div=2.0, V_init=(0.8, 1.2), div_check_interval=10)
    with pytest.raises(sc.UnsupportedError, match="wide path"):
        sc.simulate(net, sc.CLE(0.01), KIN, population=pop, cell_num=64, T=0.1, ctx=ctx)
    with pytest.raises(sc.UnsupportedError, match="wide path"):
        sc.simulate(net, sc.CLE(0.01), KIN, cell_num=64, T=0.1, precision="fp32_valu", ctx=ctx)


def test_wide_and_fused_networks_in_one_plan(ctx):
    """A plan holding two wide networks (G=300, 270) beside a fused one (G=20):
    each network's offsets into the shared pools (state, promoters, snapshots,
    injected noise) are honoured, and every network matches the oracle on its
    own injected noise.  The first network's 301 x 37 cell*genes put the
    second (40 cells, a multiple of 4) at a promoter-pool offset that is not a
    multiple of 4, so it takes the one-cell-per-lane form."""
    nets = [network(301, 40, telegraph=True), network(270, 41), sc.samplers.gpu_g20_workload()["net"]]
    cells, dt, T, seed = [37, 40, 64], 0.01, 0.3, 8
    steps = sc.n_steps(T, dt)
    rng = np.random.default_rng(seed)
    xis = [rng.standard_normal((steps, n.G, 2, c)).astype(np.float32) for n, c in zip(nets, cells)]
    print(f"wide + fused plan: G={[n.G for n in nets]} cells={cells}", flush=True)
    out = sc.simulate_gpu_batch(nets, sc.CLE(dt), KIN, cell_num=cells, T=T, readout="both", seed=seed,
                                injected_noise=xis, snapshot_every=10, full=True, ctx=ctx)
    for i, (n, c) in enumerate(zip(nets, cells)):
        ref = orc.cle(n, KIN, dt=dt, n_steps=steps, cell_num=c, injected=xis[i], seed=seed, net_id=i,
                      snapshot_every=10, threads=4)
        err = relerr(out[i].y, np.concatenate([ref["m"].T, ref["p"].T], axis=1))
        print(f"network {i}: max rel err {err:.3e}")
        assert err <= 1e-5
        np.testing.assert_array_equal(out[i].snapshots[-1, 1].T.astype(np.float64), out[i].y[:, n.G:])


def test_wide_path_large_network(ctx):
    """G=4096 (512 rows of elementwise blocks, a 64 MB FP32 W): FP32 and BF16
    against the oracle's replay of four cells on the free Philox stream."""
    G, cells, dt, T, seed = 4096, 1024, 0.01, 0.1, 3
    rng = np.random.default_rng(77)
    A = np.where(rng.random((G, G)) < 0.01, rng.uniform(0.5, 3.0, (G, G)), 0.0)
    A *= np.where(rng.random((G, G)) < 0.5, 1.0, -1.0)
    np.fill_diagonal(A, 0.0)
    net = sc.GeneNetwork(G, rng.uniform(0.5, 3.0, G), A, [], [])
    ref = orc.cle(net, KIN, dt=dt, T=T, cell_num=cells, cell_lo=0, cell_hi=4, seed=seed, threads=4)
    want = np.concatenate([ref["m"].T, ref["p"].T], axis=1)
    for precision, bar in (("fp32", 2e-4), ("bf16", 2e-2)):
        print(f"wide G={G} cells={cells} {precision}", flush=True)
        y = sc.simulate(net, sc.CLE(dt), KIN, cell_num=cells, T=T, readout="both", seed=seed, precision=precision,
                        ctx=ctx)
        err = relerr(y[:4], want)
        print(f"max rel err {err:.3e}")
        assert np.isfinite(y).all() and err <= bar, precision
