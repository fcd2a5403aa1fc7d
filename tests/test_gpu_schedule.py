"""Segmented schedule of the FP32 stepper.

When a network has more 64-cell tiles than the device holds at once, each
tile's horizon is split into segments that later workgroups continue from the
state pool.  Cells are independent and their noise is keyed by the cell index,
so the first cells of a segmented run must equal, bit for bit, a smaller run
that fits in one round (no segments); the last cells, which run in the later
rounds, are replayed by the FP64 oracle.
"""
import numpy as np
import pytest

import scrna_amd as sc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

KIN = sc.KineticParams(k_t=5.0, K_d=10.0, n=2.0, mu_m=1.0, mu_p=0.2)


def relerr(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0)))


@pytest.mark.parametrize("cells", [40000, 70001])
def test_segmented_run_matches_single_round(ctx, cells):
    # G = 64 at 4 genes per lane is a 16-wave workgroup: one per CU, so 256 tiles
    # fit at once and 625 (or 1094) tiles need segments.
    rng = np.random.default_rng(64)
    net = sc.samplers.mixed_network(64, rng, n_tf=(6, 8), targets_per_tf=(4, 10), n_and=(2, 3), n_or=(2, 3),
                                    telegraph=True)
    dt, T, seed = 0.01, 3.0, 9
    print(f"segmented schedule: G=64 cells={cells}", flush=True)
    big = sc.simulate(net, sc.CLE(dt), KIN, cell_num=cells, T=T, readout="both", seed=seed, snapshot_every=100,
                      full=True)
    small = sc.simulate(net, sc.CLE(dt), KIN, cell_num=8192, T=T, readout="both", seed=seed, snapshot_every=100,
                        full=True)
    assert np.array_equal(big.y[:8192], small.y)
    assert np.array_equal(big.snapshots[:, :, :, :8192], small.snapshots)
    lo = cells - 64
    ref = orc.cle(net, KIN, dt=dt, T=T, cell_num=cells, cell_lo=lo, cell_hi=cells, seed=seed, threads=8)
    assert relerr(big.y[lo:], np.concatenate([ref["m"].T, ref["p"].T], axis=1)) <= 2e-4
