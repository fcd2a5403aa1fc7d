"""Serving loop: requests that run concurrently on the device give exactly the
result they give alone (same seed and stream id), cancellation works, and
latency statistics are reported."""
import time

import numpy as np
import pytest

import scrna_amd as sc

pytestmark = pytest.mark.gpu

KIN = sc.KineticParams(k_t=5.0, K_d=10.0, n=2.0, mu_m=1.0, mu_p=0.2)


def make_requests():
    rng = np.random.default_rng(21)
    reqs = []
    for i, (G, cells, T, tele) in enumerate([(10, 700, 3.0, True), (50, 1500, 1.5, False), (37, 333, 4.0, True),
                                              (10, 2100, 0.5, False), (64, 129, 2.7, True), (5, 64, 0.0, False)]):
        net = sc.samplers.mixed_network(G, rng, n_tf=(2, 4), targets_per_tf=(2, 5), n_and=(0, 2), n_or=(0, 2),
                                        telegraph=tele)
        cap = sc.CaptureModel(0.1, 0.3, "mrna") if i % 2 == 0 else None
        reqs.append(dict(net=net, cells=cells, T=T, cap=cap, seed=1000 + i, net_id=i))
    return reqs


def test_mix_equals_alone(ctx):
    reqs = make_requests()
    alone = []
    for r in reqs:
        alone.append(sc.simulate(r["net"], sc.CLE(0.01), KIN, cell_num=r["cells"], T=r["T"], readout="both",
                                 capture=r["cap"], seed=r["seed"], full=True, ctx=ctx))
    with sc.Server(max_slots=8, ctx=ctx) as srv:
        tickets = []
        for r in reqs:
            tickets.append(srv.submit(r["net"], sc.CLE(0.01), KIN, cell_num=r["cells"], T=r["T"], readout="both",
                                      capture=r["cap"], seed=r["seed"]))
            time.sleep(0.002)                  # staggered arrivals join running rounds
        for i, t in enumerate(tickets):
            got = srv.wait(t)
            # the plan path uses net_id 0 for a single network, like the served request
            np.testing.assert_array_equal(got.y, alone[i].y)
            if reqs[i]["cap"] is not None:
                np.testing.assert_array_equal(got.counts, alone[i].counts)
            np.testing.assert_allclose(got.moments[0], alone[i].moments[0], rtol=1e-12)
        st = srv.stats()
        assert st["completed"] == len(reqs)
        assert st["p99_ms"] >= st["p50_ms"] > 0


def test_mix_with_segmented_and_multichunk_rounds(ctx):
    """A request whose tiles fill the device runs segmented launches; small
    requests beside it advance several chunks per round, in launches on other
    streams.  Each result still equals the request run alone."""
    rng = np.random.default_rng(33)
    reqs = [dict(G=50, cells=60000, T=3.0, tele=True), dict(G=10, cells=500, T=25.0, tele=False),
            dict(G=20, cells=3000, T=12.0, tele=True), dict(G=10, cells=70, T=7.3, tele=True)]
    for i, r in enumerate(reqs):
        r["net"] = sc.samplers.mixed_network(r["G"], rng, n_tf=(2, 4), targets_per_tf=(2, 5), n_and=(0, 2),
                                             n_or=(0, 2), telegraph=r["tele"])
        r["cap"] = sc.CaptureModel(0.1, 0.3, "mrna") if i % 2 == 0 else None
    alone = [sc.simulate(r["net"], sc.CLE(0.01), KIN, cell_num=r["cells"], T=r["T"], readout="protein",
                         capture=r["cap"], seed=77 + i, full=True, ctx=ctx) for i, r in enumerate(reqs)]
    with sc.Server(max_slots=8, ctx=ctx) as srv:
        tickets = [srv.submit(r["net"], sc.CLE(0.01), KIN, cell_num=r["cells"], T=r["T"], readout="protein",
                              capture=r["cap"], seed=77 + i) for i, r in enumerate(reqs)]
        for i, t in enumerate(tickets):
            got = srv.wait(t)
            np.testing.assert_array_equal(got.y, alone[i].y)
            if reqs[i]["cap"] is not None:
                np.testing.assert_array_equal(got.counts, alone[i].counts)
            np.testing.assert_allclose(got.moments[0], alone[i].moments[0], rtol=1e-12)


def test_cancel_and_budget(ctx):
    net = sc.samplers.mixed_network(20, np.random.default_rng(1), n_tf=(2, 3), targets_per_tf=(2, 4),
                                    n_and=(0, 1), n_or=(0, 1))
    with sc.Server(max_slots=2, ctx=ctx) as srv:
        t1 = srv.submit(net, sc.CLE(0.01), KIN, cell_num=20000, T=50.0, seed=1)
        t2 = srv.submit(net, sc.CLE(0.01), KIN, cell_num=100, T=1.0, seed=2)
        srv.cancel(t1)
        with pytest.raises(sc.ScrnaError):
            srv.wait(t1)
        r2 = srv.wait(t2)
        assert r2.y.shape == (100, 20) and np.all(np.isfinite(r2.y))
