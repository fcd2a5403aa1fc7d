"""scrna_simulate_cle_noalloc against scrna_simulate_cle through the C ABI:
bit-identical outputs on consecutive calls of decreasing size (the shared
array keeps the previous call's contents), the FP32 and BF16 paths, injected
noise against the oracle, and a clear error for a run beyond the array."""
import ctypes as C

import numpy as np
import pytest

import scrna_amd as sc
from oracle import oracle as orc
from scrna_amd import _lib

pytestmark = pytest.mark.gpu


def _call(fn_name, ctx, net, kin, *, cells, T, dt=0.01, seed=1, capture=None, snapshot_every=0, precision=0,
          injected=None, ld=None):
    lib = _lib.load()
    G = net.G
    ld = ld or cells
    cnet, keep = net.to_c()
    inj_p = None
    if injected is not None:
        inj = np.ascontiguousarray(injected, dtype=np.float32)
        keep.append(inj)
        inj_p = inj.ctypes.data_as(_lib.c_float_p)
    run = _lib.ScrnaRun(dt, T, cells, 2, 0, seed, 0, precision, snapshot_every, inj_p, 0, 0)
    n_snap = sc.n_steps(T, dt) // snapshot_every if snapshot_every else 0
    y = np.full((ld, 2 * G), -7.0, order="F")
    counts = np.full((ld, G), -7, dtype=np.int32, order="F") if capture else None
    snaps = np.empty((n_snap, 2, G, cells), dtype=np.float32) if n_snap else None
    mom = np.empty((max(n_snap, 1), G, 5))
    out = _lib.ScrnaOutput(y.ctypes.data_as(_lib.c_double_p), ld,
                           counts.ctypes.data_as(_lib.c_int32_p) if capture else None,
                           snaps.ctypes.data_as(_lib.c_float_p) if n_snap else None,
                           mom.ctypes.data_as(_lib.c_double_p), None)
    cap = capture.to_c() if capture else None
    fn = getattr(lib, fn_name)
    if fn_name == "scrna_simulate_cle":
        rc = fn(ctx.handle, C.byref(cnet), C.byref(kin.to_c()), C.byref(run), None,
                C.byref(cap) if cap else None, C.byref(out))
    else:
        rc = fn(ctx.handle, C.byref(cnet), C.byref(kin.to_c()), C.byref(run), C.byref(cap) if cap else None,
                C.byref(out))
    del keep
    return rc, dict(y=y, counts=counts, snaps=snaps, mom=mom)


def _same(ctx, net, kin, **kw):
    rc0, a = _call("scrna_simulate_cle", ctx, net, kin, **kw)
    assert rc0 == 0, _lib.load().scrna_last_error()
    rc1, b = _call("scrna_simulate_cle_noalloc", ctx, net, kin, **kw)
    assert rc1 == 0, _lib.load().scrna_last_error()
    for k in a:
        if a[k] is not None:
            np.testing.assert_array_equal(b[k], a[k], err_msg=k)
    return b


def test_noalloc_matches_simulate_on_consecutive_calls(ctx):
    kin = sc.samplers.WORKLOAD_KINETICS
    rng = np.random.default_rng(3)
    big = sc.samplers.mixed_network(20, rng, n_tf=(4, 5), targets_per_tf=(3, 6), n_and=(1, 2), n_or=(1, 2),
                                    telegraph=True)
    print("noalloc call 1: G=20, 200000 cells, snapshots, capture", flush=True)
    _same(ctx, big, kin, cells=200_000, T=1.0, seed=5, capture=sc.CaptureModel(0.1, 0.3, "mrna"),
          snapshot_every=50)
    small = sc.samplers.mixed_network(7, rng, n_tf=(2, 3), targets_per_tf=(1, 3), n_and=(0, 1), n_or=(0, 1))
    steps = sc.n_steps(0.5, 0.01)
    xi = rng.standard_normal((steps, 7, 2, 300)).astype(np.float32)
    print("noalloc call 2: G=7, 300 cells, injected noise, ld 301", flush=True)
    b = _same(ctx, small, kin, cells=300, T=0.5, seed=6, injected=xi, ld=301)
    ref = orc.cle(small, kin, dt=0.01, n_steps=steps, cell_num=300, injected=xi, seed=6)
    want = np.concatenate([ref["m"].T, ref["p"].T], axis=1)
    assert np.max(np.abs(b["y"][:300] - want) / np.maximum(np.abs(want), 1.0)) <= 1e-5
    assert np.all(b["y"][300] == -7.0)                  # rows past the cells are untouched


def test_noalloc_dense_path_and_limit(ctx):
    kin = sc.samplers.WORKLOAD_KINETICS
    net = sc.samplers.mixed_network(64, np.random.default_rng(4), n_tf=(8, 10), targets_per_tf=(3, 6),
                                    n_and=(1, 2), n_or=(1, 2))
    print("noalloc dense BF16 G=64", flush=True)
    _same(ctx, net, kin, cells=4096, T=0.5, seed=2, precision=1)
    lib = _lib.load()
    limit = lib.scrna_noalloc_scratch_bytes()
    assert limit == 128 << 20
    cells = int(limit // (9 * 64)) + 1000               # state and promoters alone exceed the array
    print(f"noalloc over the limit: G=64, {cells} cells", flush=True)
    rc, _ = _call("scrna_simulate_cle_noalloc", ctx, net, kin, cells=cells, T=0.01)
    assert rc == _lib.SCRNA_ERR_INVALID
    assert b"scratch" in lib.scrna_last_error()
    _same(ctx, net, kin, cells=100, T=0.3, seed=9)        # still usable after a rejected call
