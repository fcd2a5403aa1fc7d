"""Seeded sweep of random small configurations against the FP64 oracle on
identical injected noise: network size and shape, kinetics (Hill n in
{1, 1.7, 2, 3}), telegraph or constitutive promoters, cells per tile (ragged
tiles, single cells), step size and horizon, and the stepper's genes-per-lane
tile (sparse FP32 stepper) or the dense forms (FP32 on the matrix cores,
BF16x3).  Bars as in test_gpu_parity / test_gpu_dense: 1e-5 relative for FP32,
2e-4 for BF16x3, relative to max(|x|, 1).
"""
import numpy as np
import pytest

import scrna_amd as sc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0)))


def random_kinetics(rng):
    return sc.KineticParams(k_t=float(rng.uniform(1.0, 10.0)), K_d=float(rng.uniform(2.0, 20.0)),
                            n=float(rng.choice([1.0, 1.7, 2.0, 3.0])), mu_m=float(rng.uniform(0.5, 2.0)),
                            mu_p=float(rng.uniform(0.1, 0.5)))


def compare(net, kin, cells, dt, T, seed, precision, ctx):
    steps = sc.n_steps(T, dt)
    rng = np.random.default_rng(seed)
    xi = rng.standard_normal((steps, net.G, 2, cells)).astype(np.float32)
    y = sc.simulate(net, sc.CLE(dt), kin, cell_num=cells, T=T, readout="both", injected_noise=xi,
                    seed=seed, precision=precision, ctx=ctx)
    ref = orc.cle(net, kin, dt=dt, n_steps=steps, cell_num=cells, injected=xi, seed=seed)
    return relerr(y, np.concatenate([ref["m"].T, ref["p"].T], axis=1))


@pytest.mark.parametrize("case", range(16))
def test_sparse_stepper_sweep(case):
    rng = np.random.default_rng(7100 + case)
    G = int(rng.integers(1, 65))
    net = sc.samplers.mixed_network(G, rng, n_tf=(0, min(G, 6)), targets_per_tf=(1, 8), n_and=(0, 2),
                                    n_or=(0, 2), telegraph=bool(rng.random() < 0.5))
    kin = random_kinetics(rng)
    cells = int(rng.choice([1, int(rng.integers(2, 64)), int(rng.integers(65, 200))]))
    dt = float(rng.choice([0.01, 0.005]))
    steps = int(rng.integers(20, 120))
    ch = int(rng.choice([2, 4, 8, 16]))
    ctx = sc.Context(0)
    try:
        ctx.set_tile(1, 16 * ch, ch)          # at most 16 waves per workgroup
        print(f"sweep {case}: G={G} cells={cells} dt={dt} steps={steps} tile={ch} n={kin.n} "
              f"telegraph={bool(np.isfinite(np.asarray(net.k_on, dtype=float)).any())}", flush=True)
        err = compare(net, kin, cells, dt, steps * dt, 11 + case, "fp32_valu", ctx)
    finally:
        ctx.close()
    print(f"max rel err {err:.3e}")
    assert err <= 1e-5


def random_dense_network(G, rng, density, telegraph):
    A = np.where(rng.random((G, G)) < density, rng.uniform(0.5, 3.0, (G, G)), 0.0)
    A *= np.where(rng.random((G, G)) < 0.5, 1.0, -1.0)
    np.fill_diagonal(A, 0.0)
    coop, redun = [], []
    for kind in ("and", "or"):
        for _ in range(int(rng.integers(0, 3))):
            t = int(rng.integers(0, G))
            src = sorted(int(x) for x in rng.choice([j for j in range(G) if j != t], size=2, replace=False))
            (coop if kind == "and" else redun).append(
                (sc.CoopEdge if kind == "and" else sc.RedunEdge)(t, src, float(rng.uniform(2.0, 5.0))))
    k_on = rng.uniform(0.1, 1.0, G) if telegraph else None
    k_off = rng.uniform(0.5, 5.0, G) if telegraph else None
    return sc.GeneNetwork(G, rng.uniform(0.5, 3.0, G), A, coop, redun, k_on=k_on, k_off=k_off)


@pytest.mark.parametrize("case", range(8))
def test_dense_forms_sweep(ctx, case):
    rng = np.random.default_rng(7300 + case)
    G = int(rng.integers(20, 257))
    net = random_dense_network(G, rng, float(rng.uniform(0.05, 0.4)), bool(rng.random() < 0.3))
    kin = random_kinetics(rng)
    cells = int(rng.integers(1, 80))
    dt = float(rng.choice([0.01, 0.005]))
    steps = int(rng.integers(10, 60))
    for precision, bar in (("fp32_mfma", 1e-5), ("bf16x3", 2e-4)):
        print(f"dense sweep {case}: G={G} cells={cells} dt={dt} steps={steps} n={kin.n} {precision}", flush=True)
        err = compare(net, kin, cells, dt, steps * dt, 31 + case, precision, ctx)
        print(f"max rel err {err:.3e}")
        assert err <= bar, precision
