"""GPU simulation entry points (the drop-in for ext/SyntheticscRNAseqCUDAExt.jl).

``simulate(net, CLE(dt), kin, "gpu"; cell_num, T, readout)`` replaces the
CUDA.jl method at ext:48-117 (and ext:409-529 with a PopulationConfig);
``simulate(net, BinomialTauLeap(dt), kin, "gpu"; ...)`` replaces ext:230-311;
``simulate_gpu_batch`` replaces ext:132-213 and ext:317-399; ``apply_capture``
is the GPU form of src/capture.jl:25-47.  All work runs in libscrna.so's HIP
kernels.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from . import _lib
from .kinetics import CaptureModel, KineticParams, PopulationConfig
from .network import GeneNetwork


@dataclass(frozen=True)
class CLE:
    """Euler-Maruyama CLE with collapsed noise (cle.jl:21-23)."""
    dt: float


@dataclass(frozen=True)
class BinomialTauLeap:
    """Binomial tau-leap (tauleap_binomial.jl:19-22): Poisson creation and
    binomial decay per step, integer counts.  The GPU draws are exact (the
    reference GPU form truncates a Poisson inverse CDF and approximates the
    binomial decay by a clamped Poisson, ext:21-39)."""
    dt: float


def _algorithm(alg) -> int:
    if isinstance(alg, CLE):
        return 0
    if isinstance(alg, BinomialTauLeap):
        return 1
    raise TypeError("the GPU path implements CLE and BinomialTauLeap")


@dataclass
class SimResult:
    y: np.ndarray                      # readout, cells x G (x 2G for "both")
    counts: np.ndarray | None          # captured counts, cells x G int32
    snapshots: np.ndarray | None       # (n_snap, 2, G, cells) float32; species 0 = mRNA, 1 = protein
    moments: np.ndarray | None         # (n_snap or 1, G, 5) sums over cells: m, m^2, p, p^2, m*p
    volumes: np.ndarray | None         # (cells,) with population


def n_steps(T: float, dt: float) -> int:
    """ceil(T/dt) in double (cle.jl:56, ext:69)."""
    return int(np.ceil(T / dt))


def _seed(seed) -> int:
    if seed is None:
        return int.from_bytes(os.urandom(8), "little")
    return int(seed) & 0xFFFFFFFFFFFFFFFF


def _run(alg, cell_num: int, T: float, readout: str, init: str, seed, net_id: int,
         precision: str, snapshot_every: int, injected) -> tuple:
    if readout not in _lib.READOUT:
        raise ValueError("readout must be 'mrna', 'protein' or 'both'")
    if init not in ("steady", "zeros"):
        raise ValueError("init must be 'steady' (the steady state, cle.jl:52-53 / tauleap_binomial.jl:61-62) "
                         "or 'zeros'")
    if precision not in _lib.PRECISION:
        raise ValueError("precision must be 'fp32', 'fp32_valu', 'fp32_mfma', 'bf16', 'bf16x3' or 'fp32_unfused'")
    keep = []
    inj_p = None
    if injected is not None:
        inj = np.ascontiguousarray(injected, dtype=np.float32)
        keep.append(inj)
        inj_p = inj.ctypes.data_as(_lib.c_float_p)
    r = _lib.ScrnaRun(float(alg.dt), float(T), int(cell_num), _lib.READOUT[readout],
                      0 if init == "steady" else 1, _seed(seed), int(net_id), _lib.PRECISION[precision],
                      int(snapshot_every), inj_p, _algorithm(alg), 0)
    return r, keep


def _outputs(G: int, cells: int, readout: str, capture: bool, n_snap: int, n_mom: int, population: bool):
    ncol = 2 * G if readout == "both" else G
    y = np.empty((cells, ncol), dtype=np.float64, order="F")
    counts = np.empty((cells, G), dtype=np.int32, order="F") if capture else None
    snaps = np.empty((n_snap, 2, G, cells), dtype=np.float32) if n_snap > 0 else None
    mom = np.empty((n_mom, G, 5), dtype=np.float64)
    vol = np.empty(cells, dtype=np.float64) if population else None
    out = _lib.ScrnaOutput(
        y.ctypes.data_as(_lib.c_double_p), cells,
        counts.ctypes.data_as(_lib.c_int32_p) if counts is not None else None,
        snaps.ctypes.data_as(_lib.c_float_p) if snaps is not None else None,
        mom.ctypes.data_as(_lib.c_double_p),
        vol.ctypes.data_as(_lib.c_double_p) if vol is not None else None)
    return out, SimResult(y, counts, snaps, mom, vol)


def _prepare(nets, kins, runs_kw, pops, caps):
    n = len(nets)
    NetA = _lib.ScrnaNetwork * n
    KinA = _lib.ScrnaKinetics * n
    RunA = _lib.ScrnaRun * n
    PopA = _lib.ScrnaPopulation * n
    CapA = _lib.ScrnaCapture * n
    keep = []
    cn, ck, cr, cp, cc = NetA(), KinA(), RunA(), PopA(), CapA()
    any_pop = any(p is not None for p in pops)
    any_cap = any(c is not None for c in caps)
    for i in range(n):
        s, k = nets[i].to_c()
        keep.append(k)
        cn[i] = s
        ck[i] = kins[i].to_c()
        r, k2 = _run(**runs_kw[i])
        keep.append(k2)
        cr[i] = r
        if pops[i] is not None:
            cp[i] = pops[i].to_c()
        if caps[i] is not None:
            cc[i] = caps[i].to_c()
    return cn, ck, cr, (cp if any_pop else None), (cc if any_cap else None), keep


def _listify(x, n):
    if isinstance(x, (list, tuple)):
        if len(x) != n:
            raise ValueError("per-network argument lists must match the number of networks")
        return list(x)
    return [x] * n


def simulate_gpu_batch(networks: Sequence[GeneNetwork], alg, kin, *, cell_num=1000, T=50.0,
                       readout: str = "protein", population=None, capture=None, seed=None,
                       init: str = "steady", snapshot_every: int = 0, precision: str = "fp32",
                       injected_noise=None, full: bool = False, ctx: _lib.Context | None = None):
    """Batch-simulate networks in one plan (ext:132-213).

    Unlike the reference, networks may differ in G; ``kin``, ``cell_num``,
    ``T``, ``population``, ``capture`` and ``injected_noise`` may be given per
    network as lists.  Network i draws the Philox stream (seed, net_id=i).
    Returns a list of readout matrices (or SimResult objects with full=True).
    """
    nets = list(networks)
    n = len(nets)
    if n == 0:
        raise ValueError("no networks")
    _algorithm(alg)
    ctx = ctx or _lib.default_context()
    kins = _listify(kin, n)
    cells = _listify(cell_num, n)
    Ts = _listify(T, n)
    pops = _listify(population, n)
    caps = _listify(capture, n)
    injs = _listify(injected_noise, n) if isinstance(injected_noise, (list, tuple)) else [injected_noise] * n
    s = _seed(seed)
    runs_kw = []
    for i in range(n):
        c = pops[i].cell_num if pops[i] is not None else cells[i]
        runs_kw.append(dict(alg=alg, cell_num=c, T=Ts[i], readout=readout, init=init, seed=s, net_id=i,
                            precision=precision, snapshot_every=snapshot_every, injected=injs[i]))
    cn, ck, cr, cp, cc, keep = _prepare(nets, kins, runs_kw, pops, caps)
    OutA = _lib.ScrnaOutput * n
    outs = OutA()
    results = []
    for i in range(n):
        ns = n_steps(Ts[i], alg.dt)
        n_snap = ns // snapshot_every if snapshot_every > 0 else 0
        o, res = _outputs(nets[i].G, cr[i].cell_num, readout, caps[i] is not None, n_snap, max(n_snap, 1),
                          pops[i] is not None)
        outs[i] = o
        results.append(res)
    lib = _lib.load()
    _lib.check(lib.scrna_simulate_cle_batch(ctx.handle, n, cn, ck, cr, cp, cc, outs), "simulate_gpu_batch")
    del keep
    return results if full else [r.y for r in results]


def simulate(net: GeneNetwork, alg, kin: KineticParams, backend: str = "gpu",
             population: PopulationConfig | None = None, *, cell_num: int = 1000, T: float = 50.0,
             readout: str = "protein", capture: CaptureModel | None = None, seed=None,
             init: str = "steady", snapshot_every: int = 0, precision: str = "fp32",
             injected_noise=None, full: bool = False, ctx: _lib.Context | None = None):
    """simulate(net, CLE(dt), kin, "gpu"[, pop]; cell_num, T, readout) -> cells x G matrix.

    Drop-in for the CUDA extension methods (ext:48-52 and ext:409-412).  The
    default readout is "protein" (cle.jl:37, ext:52).  Keyword extensions:
    ``capture`` fuses the LogNormal-Binomial capture on the final state and
    makes the call return the captured counts; ``seed`` keys the Philox
    stream; ``init="zeros"`` starts from zero counts; ``snapshot_every`` keeps
    the state every k steps; ``precision`` picks the regulation arithmetic;
    ``full=True`` returns a SimResult with everything.
    """
    if backend != "gpu":
        raise ValueError("this package implements the GPU path only (backend='gpu')")
    res = simulate_gpu_batch([net], alg, kin, cell_num=cell_num, T=T, readout=readout,
                             population=population, capture=capture, seed=seed, init=init,
                             snapshot_every=snapshot_every, precision=precision,
                             injected_noise=[injected_noise], full=True, ctx=ctx)[0]
    if full:
        return res
    if capture is not None:
        return res.counts
    return res.y


def apply_capture(X, capture: CaptureModel, *, seed=None, net_id: int = 0,
                  ctx: _lib.Context | None = None) -> np.ndarray:
    """GPU LogNormal-Binomial capture of a cells x G count matrix (capture.jl:25-47)."""
    ctx = ctx or _lib.default_context()
    X = np.asfortranarray(X, dtype=np.float64)
    if X.ndim != 2:
        raise ValueError("X must be a cells x G matrix")
    cells, G = X.shape
    Y = np.empty((cells, G), dtype=np.int32, order="F")
    lib = _lib.load()
    _lib.check(lib.scrna_apply_capture(ctx.handle, X.ctypes.data_as(_lib.c_double_p), cells, G, cells,
                                       capture.efficiency, capture.efficiency_std, _seed(seed), net_id,
                                       Y.ctypes.data_as(_lib.c_int32_p)), "apply_capture")
    return Y


class Plan:
    """Device-resident batch: upload once, launch many times (scrna_plan_*).

    Arguments as simulate_gpu_batch; network i draws the stream (seed, net_id=i).
    The plan copies ``injected_noise`` at creation."""

    def __init__(self, networks, alg, kin, *, cell_num=1000, T=50.0, readout="protein",
                 population=None, capture=None, seed=None, init="steady", snapshot_every=0,
                 precision="fp32", injected_noise=None, ctx: _lib.Context | None = None):
        nets = list(networks)
        n = len(nets)
        self.ctx = ctx or _lib.default_context()
        self.lib = _lib.load()
        kins = _listify(kin, n)
        cells = _listify(cell_num, n)
        Ts = _listify(T, n)
        pops = _listify(population, n)
        caps = _listify(capture, n)
        injs = _listify(injected_noise, n) if isinstance(injected_noise, (list, tuple)) else [injected_noise] * n
        s = _seed(seed)
        runs_kw = [dict(alg=alg, cell_num=(pops[i].cell_num if pops[i] is not None else cells[i]), T=Ts[i],
                        readout=readout, init=init, seed=s, net_id=i, precision=precision,
                        snapshot_every=snapshot_every, injected=injs[i]) for i in range(n)]
        cn, ck, cr, cp, cc, keep = _prepare(nets, kins, runs_kw, pops, caps)
        h = C.c_void_p()
        _lib.check(self.lib.scrna_plan_create(self.ctx.handle, n, cn, ck, cr, cp, cc, C.byref(h)),
                   "scrna_plan_create")
        self.handle = h
        self.nets, self.runs, self.pops, self.caps = nets, [cr[i] for i in range(n)], pops, caps
        self.readout, self.snapshot_every, self.alg = readout, snapshot_every, alg
        self.Ts = Ts

    def launch(self, stream: int | None = None) -> None:
        _lib.check(self.lib.scrna_plan_launch(self.handle, C.c_void_p(stream) if stream else None),
                   "scrna_plan_launch")

    def sync(self) -> None:
        _lib.check(self.lib.scrna_plan_sync(self.handle), "scrna_plan_sync")

    def use_graph(self) -> None:
        """Capture the whole simulation into one hipGraph replayed by launch()."""
        _lib.check(self.lib.scrna_plan_use_graph(self.handle), "scrna_plan_use_graph")

    def stepper_ms(self) -> tuple:
        ms, cgs = C.c_double(), C.c_double()
        _lib.check(self.lib.scrna_plan_stepper_ms(self.handle, C.byref(ms), C.byref(cgs)), "stepper_ms")
        return ms.value, cgs.value

    def nbytes(self) -> int:
        b = C.c_uint64()
        _lib.check(self.lib.scrna_plan_bytes(self.handle, C.byref(b)), "scrna_plan_bytes")
        return b.value

    def download(self, i: int = 0) -> SimResult:
        ns = n_steps(self.Ts[i], self.alg.dt)
        n_snap = ns // self.snapshot_every if self.snapshot_every > 0 else 0
        o, res = _outputs(self.nets[i].G, self.runs[i].cell_num, self.readout, self.caps[i] is not None,
                          n_snap, max(n_snap, 1), self.pops[i] is not None)
        _lib.check(self.lib.scrna_plan_download(self.handle, i, C.byref(o)), "scrna_plan_download")
        return res

    def close(self) -> None:
        if self.handle:
            self.lib.scrna_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
