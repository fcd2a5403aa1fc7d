"""The benchmark workloads' own configurations against the oracle, and the C
entry point the Julia shim binds, called directly through ctypes.

* The CPU reference workload (G=5 mixed network, constitutive, 5000 cells
  from zero counts, dt=0.01, T=50, LogNormal(0.1, 0.3)-Binomial capture) at
  full size: every cell against the FP64 oracle replaying the same Philox
  stream (5e-5: GPU FP32 normals against the FP64 transform of the same
  u32), the fused capture bit-exact against the oracle's capture of the GPU
  state, and the per-gene means against the oracle's within 0.5%.
* The serving workload's request mix (G in {10, 50, 200}, half telegraph,
  capture, horizons up to T=200 at dt=0.01) at reduced cell counts with
  Poisson arrivals: each served result equals the same request alone, and
  cells of every request replay against the oracle.
* scrna_simulate_cle (SyntheticscRNAseqHIP.jl's binding) with hand-built
  structs: injected-noise parity, capture, a population run and the error
  path.
"""
import ctypes as C
import time

import numpy as np
import pytest

import scrna_amd as sc
from oracle import oracle as orc
from scrna_amd import _lib

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0)))


def test_cpu_reference_workload_full_size(ctx):
    w = sc.samplers.cpu_g5_workload()
    net, kin = w["net"], w["kin"]
    G, cells, dt, T = net.G, w["cells"], w["dt"], w["T"]
    print(f"G=5 workload: {cells} cells, {sc.n_steps(T, dt)} steps, zero init, capture", flush=True)
    res = sc.simulate(net, sc.CLE(dt), kin, cell_num=cells, T=T, readout="both", init=w["init"], capture=w["capture"],
                      seed=2024, full=True)
    ref = orc.cle(net, kin, dt=dt, T=T, cell_num=cells, seed=2024, init="zeros", threads=8)
    want = np.concatenate([ref["m"].T, ref["p"].T], axis=1)
    err = relerr(res.y, want)
    print(f"max rel err over {cells} cells: {err:.2e}")
    assert err <= 5e-5
    Yo, _ = orc.capture(res.y[:, :G].T, w["capture"].efficiency, w["capture"].efficiency_std, 2024)
    np.testing.assert_array_equal(res.counts, Yo.T)
    np.testing.assert_allclose(res.y.mean(axis=0), want.mean(axis=0), rtol=5e-3)
    # E[Y] = efficiency * E[X] (capture.jl:30: mean-corrected LogNormal)
    ratio = res.counts.mean(axis=0) / np.maximum(res.y[:, :G].mean(axis=0), 1e-9)
    assert np.all(np.abs(ratio - 0.1) < 0.01)


def serving_requests(seed=5):
    rng = np.random.default_rng(seed)
    kin = sc.samplers.WORKLOAD_KINETICS
    specs = [(200, 1500, 200.0, True), (10, 4000, 10.0, False), (50, 2200, 120.0, True), (200, 1000, 37.5, False),
             (10, 1200, 200.0, True), (50, 3000, 64.0, False)]
    reqs = []
    for i, (G, cells, T, tele) in enumerate(specs):
        net = sc.samplers.mixed_network(G, rng, n_tf=(2, 3) if G <= 10 else (4, 5),
                                        targets_per_tf=(2, 4) if G <= 10 else (3, 6),
                                        n_and=(0, 1) if G <= 10 else (1, 2), n_or=(0, 1) if G <= 10 else (1, 2),
                                        telegraph=tele)
        reqs.append(dict(net=net, kin=kin, cells=cells, T=T, seed=1000 + i))
    return reqs


def test_serving_workload_mix_against_oracle(ctx):
    reqs = serving_requests()
    cap = sc.CaptureModel(0.1, 0.3, "mrna")
    alone = [sc.simulate(r["net"], sc.CLE(0.01), r["kin"], cell_num=r["cells"], T=r["T"], readout="both",
                         capture=cap, seed=r["seed"], full=True) for r in reqs]
    gaps = np.random.default_rng(9).exponential(1.0 / 8.0, size=len(reqs))     # Poisson arrivals, 8 req/s
    with sc.Server(max_slots=16) as srv:
        tickets = []
        for r, gap in zip(reqs, gaps):
            time.sleep(gap)
            tickets.append(srv.submit(r["net"], sc.CLE(0.01), r["kin"], cell_num=r["cells"], T=r["T"],
                                      readout="both", capture=cap, seed=r["seed"]))
        got = [srv.wait(t) for t in tickets]
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["p99_ms"] >= st["p50_ms"] > 0
    for i, r in enumerate(reqs):
        np.testing.assert_array_equal(got[i].y, alone[i].y)
        np.testing.assert_array_equal(got[i].counts, alone[i].counts)
        G, cells = r["net"].G, r["cells"]
        # first two and last cell of the request, replayed in FP64 from the same Philox stream
        for lo, hi in ((0, 2), (cells - 1, cells)):
            ref = orc.cle(r["net"], r["kin"], dt=0.01, T=r["T"], cell_num=cells, cell_lo=lo, cell_hi=hi,
                          seed=r["seed"], threads=2)
            want = np.concatenate([ref["m"].T, ref["p"].T], axis=1)
            err = relerr(got[i].y[lo:hi], want)
            print(f"request {i} G={G} T={r['T']} cells [{lo},{hi}): max rel err {err:.2e}", flush=True)
            assert err <= 5e-5
        Yo, _ = orc.capture(got[i].y[:, :G].T, 0.1, 0.3, r["seed"])
        np.testing.assert_array_equal(got[i].counts, Yo.T)


def _c_structs(net, kin, dt, T, cells, *, seed, readout=2, injected=None, init=0):
    cnet, keep = net.to_c()
    inj_p = None
    if injected is not None:
        inj = np.ascontiguousarray(injected, dtype=np.float32)
        keep.append(inj)
        inj_p = inj.ctypes.data_as(_lib.c_float_p)
    run = _lib.ScrnaRun(dt, T, cells, readout, init, seed, 0, 0, 0, inj_p, 0, 0)
    return cnet, kin.to_c(), run, keep


def test_simulate_cle_entry_point_direct(ctx):
    lib = _lib.load()
    rng = np.random.default_rng(17)
    net = sc.samplers.mixed_network(12, rng, n_tf=(3, 4), targets_per_tf=(2, 4), n_and=(1, 1), n_or=(1, 1),
                                    telegraph=True)
    kin = sc.samplers.WORKLOAD_KINETICS
    cells, dt, T = 333, 0.01, 1.5
    steps = sc.n_steps(T, dt)
    xi = rng.standard_normal((steps, net.G, 2, cells)).astype(np.float32)
    cnet, ckin, run, keep = _c_structs(net, kin, dt, T, cells, seed=3, injected=xi)
    cap = _lib.ScrnaCapture(1, 0.1, 0.3, 0)
    y = np.empty((cells, 2 * net.G), order="F")
    counts = np.empty((cells, net.G), dtype=np.int32, order="F")
    mom = np.empty((1, net.G, 5))
    out = _lib.ScrnaOutput(y.ctypes.data_as(_lib.c_double_p), cells, counts.ctypes.data_as(_lib.c_int32_p), None,
                           mom.ctypes.data_as(_lib.c_double_p), None)
    rc = lib.scrna_simulate_cle(ctx.handle, C.byref(cnet), C.byref(ckin), C.byref(run), None, C.byref(cap),
                                C.byref(out))
    assert rc == 0, lib.scrna_last_error()
    ref = orc.cle(net, kin, dt=dt, n_steps=steps, cell_num=cells, injected=xi, seed=3)
    assert relerr(y, np.concatenate([ref["m"].T, ref["p"].T], axis=1)) <= 1e-5
    Yo, _ = orc.capture(y[:, :net.G].T, 0.1, 0.3, 3)
    np.testing.assert_array_equal(counts, Yo.T)
    np.testing.assert_allclose(mom[0], orc.moments(y[:, :net.G].T, y[:, net.G:].T), rtol=1e-6)
    # the same call through the batch entry gives the same bits
    res = sc.simulate(net, sc.CLE(dt), kin, cell_num=cells, T=T, readout="both", injected_noise=xi, seed=3,
                      capture=sc.CaptureModel(0.1, 0.3, "mrna"), full=True)
    np.testing.assert_array_equal(res.y, y)
    del keep


def test_simulate_cle_entry_point_population_and_errors(ctx):
    lib = _lib.load()
    net = sc.GeneNetwork(2, np.array([3.0, 2.0]), np.array([[0, 1.5], [-1.0, 0]]))
    kin = sc.KineticParams(k_t=2.0, K_d=10.0, n=2.0, mu_m=0.5, mu_p=0.2)
    cells, dt, T = 200, 0.05, 4.0
    steps = sc.n_steps(T, dt)
    xi = np.random.default_rng(2).standard_normal((steps, 2, 2, cells)).astype(np.float32)
    cnet, ckin, run, keep = _c_structs(net, kin, dt, T, cells, seed=8, injected=xi)
    pcfg = sc.PopulationConfig(cell_num=cells, growth_rate=0.15, V_div=2.0, V_init=(0.9, 1.9), div_check_interval=5,
                               partition="halve", replace="exclude_mother")
    pop = pcfg.to_c()
    y = np.empty((cells, 4), order="F")
    vol = np.empty(cells)
    out = _lib.ScrnaOutput(y.ctypes.data_as(_lib.c_double_p), cells, None, None, None,
                           vol.ctypes.data_as(_lib.c_double_p))
    rc = lib.scrna_simulate_cle(ctx.handle, C.byref(cnet), C.byref(ckin), C.byref(run), C.byref(pop), None,
                                C.byref(out))
    assert rc == 0, lib.scrna_last_error()
    ref = orc.cle_population(net, kin, pcfg, dt=dt, T=T, cell_num=cells, seed=8, injected=xi)
    np.testing.assert_array_equal(vol, ref["V"])
    assert relerr(y, np.concatenate([ref["m"].T, ref["p"].T], axis=1)) <= 1e-4
    # network.jl:68: a nonzero diagonal is an AssertionError -> SCRNA_ERR_INVALID with a message
    bad = sc.GeneNetwork.__new__(sc.GeneNetwork)
    bad.__dict__.update(net.__dict__)
    bad.interactions = np.array([[1.0, 1.5], [-1.0, 0]])
    bnet, bkeep = bad.to_c()
    rc = lib.scrna_simulate_cle(ctx.handle, C.byref(bnet), C.byref(ckin), C.byref(run), None, None, C.byref(out))
    assert rc == _lib.SCRNA_ERR_INVALID
    assert b"diagonal" in lib.scrna_last_error()
    del keep, bkeep


def test_captured_counts_to_10x_directory(ctx, tmp_path):
    """The GPU path end to end into the on-disk format the reference's real-data
    comparison reads (experiments/compare_real_scrnaseq.jl:62-96): simulate with
    fused capture, write the 10x v2 directory, read it back (genes x cells)."""
    import scipy.io

    rng = np.random.default_rng(41)
    net = sc.samplers.mixed_network(20, rng, n_tf=(4, 5), targets_per_tf=(3, 6), n_and=(1, 2), n_or=(1, 2),
                                    telegraph=True)
    res = sc.simulate(net, sc.CLE(0.01), sc.samplers.WORKLOAD_KINETICS, cell_num=3000, T=5.0, readout="mrna",
                      capture=sc.CaptureModel(0.1, 0.3, "mrna"), seed=12, full=True)
    nnz = sc.io.write_10x(str(tmp_path / "sim"), res.counts, barcode_prefix="SIM")
    M = scipy.io.mmread(str(tmp_path / "sim" / "matrix.mtx")).toarray()
    np.testing.assert_array_equal(M, res.counts.T)
    assert nnz == np.count_nonzero(res.counts)
    genes = (tmp_path / "sim" / "genes.tsv").read_text().splitlines()
    barcodes = (tmp_path / "sim" / "barcodes.tsv").read_text().splitlines()
    assert len(genes) == 20 and len(barcodes) == 3000
