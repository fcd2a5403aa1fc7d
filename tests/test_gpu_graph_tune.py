"""hipGraph replay and the tile-shape autotuner.

A captured plan is replayed several times with eager work and
synchronisations in between; every replay must equal the eager result.
Tile shapes change the work mapping, never the arithmetic: results with any
genes-per-lane choice are identical.
"""
import json

import numpy as np
import pytest

import scrna_amd as sc

pytestmark = pytest.mark.gpu

KIN = sc.KineticParams(k_t=5.0, K_d=10.0, n=2.0, mu_m=1.0, mu_p=0.2)


def nets(seed=4):
    rng = np.random.default_rng(seed)
    return [sc.samplers.mixed_network(G, rng, n_tf=(2, 4), targets_per_tf=(2, 5), n_and=(0, 2), n_or=(0, 2),
                                      telegraph=True) for G in (7, 20, 33, 64)]


def test_graph_replays_equal_eager(ctx):
    ns = nets()
    cap = sc.CaptureModel(0.1, 0.3, "mrna")
    eager = sc.Plan(ns, sc.CLE(0.01), KIN, cell_num=[300, 1000, 77, 129], T=1.0, readout="both", capture=cap,
                    seed=5, snapshot_every=20, ctx=ctx)
    eager.launch()
    eager.sync()
    want = [eager.download(i) for i in range(len(ns))]
    graph = sc.Plan(ns, sc.CLE(0.01), KIN, cell_num=[300, 1000, 77, 129], T=1.0, readout="both", capture=cap,
                    seed=5, snapshot_every=20, ctx=ctx)
    graph.use_graph()
    for rep in range(3):
        graph.launch()
        graph.sync()
        eager.launch()            # eager work between replays
        eager.sync()
        for i in range(len(ns)):
            got = graph.download(i)
            np.testing.assert_array_equal(got.y, want[i].y)
            np.testing.assert_array_equal(got.counts, want[i].counts)
            np.testing.assert_array_equal(got.snapshots, want[i].snapshots)
            np.testing.assert_array_equal(got.moments, want[i].moments)


def test_graph_population_replay(ctx):
    net = nets()[1]
    pop = sc.PopulationConfig(cell_num=640, growth_rate=0.05, V_div=2.0, V_init=(0.8, 1.2), div_check_interval=10,
                              partition="binomial")
    eager = sc.Plan([net], sc.CLE(0.01), KIN, population=pop, T=2.0, readout="both", seed=6, ctx=ctx)
    eager.launch()
    eager.sync()
    want = eager.download(0)
    g = sc.Plan([net], sc.CLE(0.01), KIN, population=pop, T=2.0, readout="both", seed=6, ctx=ctx)
    g.use_graph()
    for rep in range(3):
        g.launch()
        g.sync()
        got = g.download(0)
        np.testing.assert_array_equal(got.y, want.y)
        np.testing.assert_array_equal(got.volumes, want.volumes)


@pytest.mark.parametrize("ch", [2, 4, 8, 16])
def test_tile_shape_does_not_change_results(ch):
    ctx2 = sc.Context(0)
    try:
        ns = nets(9)
        ref = sc.simulate_gpu_batch(ns, sc.CLE(0.01), KIN, cell_num=200, T=0.5, readout="both", seed=2, ctx=ctx2)
        ctx2.set_tile(1, min(64, 16 * ch), ch)     # at most 16 waves per workgroup
        got = sc.simulate_gpu_batch(ns, sc.CLE(0.01), KIN, cell_num=200, T=0.5, readout="both", seed=2, ctx=ctx2)
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b)
    finally:
        ctx2.close()


def test_autotune_writes_cache(tmp_path):
    ctx2 = sc.Context(0)
    try:
        path = str(tmp_path / "tiles.json")
        n = ctx2.autotune(max_G=64, cache_path=path)
        assert n >= 4
        table = json.load(open(path))["table"]
        assert table[0][0] == 1 and table[-1][1] >= 64
        assert all(t[2] in (2, 4, 8, 16) for t in table)
        assert ctx2.autotune(max_G=64, cache_path=path) == n      # cache hit
    finally:
        ctx2.close()


def test_autotune_phases_keys_its_cache_by_launch_length(tmp_path):
    """A table tuned for population phases is cached under its launch length and
    is not taken for whole-horizon tuning (nor the other way round)."""
    ctx2 = sc.Context(0)
    try:
        path = str(tmp_path / "tiles.json")
        n = ctx2.autotune(max_G=16, cache_path=path, phase_steps=10)
        saved = json.load(open(path))
        assert saved["arch"].endswith("/phase10")
        assert all(t[2] in (2, 4, 8, 16) for t in saved["table"]) and saved["table"][-1][1] >= 16
        assert ctx2.autotune(max_G=16, cache_path=path, phase_steps=10) == n    # cache hit
        ctx2.autotune(max_G=16, cache_path=path)                                 # different key: re-tuned
        assert not json.load(open(path))["arch"].endswith("/phase10")
    finally:
        ctx2.close()
