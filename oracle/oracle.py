"""ctypes wrapper of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module.  It is the checker, never the thing
measured or shipped.  See scrna_oracle.h for the reference files restated.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_fp = C.POINTER(C.c_float)
_up = C.POINTER(C.c_uint8)


class OrcNetwork(C.Structure):
    _fields_ = [("G", C.c_int32), ("basals", _dp), ("A", _dp),
                ("n_coop", C.c_int32), ("coop_target", _ip), ("coop_off", _ip), ("coop_src", _ip),
                ("coop_strength", _dp),
                ("n_redun", C.c_int32), ("redun_target", _ip), ("redun_off", _ip), ("redun_src", _ip),
                ("redun_strength", _dp),
                ("k_on", _dp), ("k_off", _dp)]


class OrcKinetics(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("k_t", "K_d", "n", "mu_m", "mu_p", "dilution")]


class OrcPopulation(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("growth_rate", C.c_double), ("V_div", C.c_double),
                ("V_init_lo", C.c_double), ("V_init_hi", C.c_double), ("div_check_interval", C.c_int32),
                ("partition", C.c_int32), ("replace_mode", C.c_int32), ("growth_noise", C.c_double)]


class OrcRun(C.Structure):
    _fields_ = [("dt", C.c_double), ("n_steps", C.c_int64), ("cell_num", C.c_int64),
                ("cell_lo", C.c_int64), ("cell_hi", C.c_int64), ("seed", C.c_uint64),
                ("net_id", C.c_uint32), ("init_mode", C.c_int32), ("snapshot_every", C.c_int64),
                ("injected", _fp), ("n_threads", C.c_int32)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        lib.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        lib.orc_normal_pair.argtypes = [C.c_uint32, C.c_uint32, _dp, _dp]
        lib.orc_cle_flip.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32,
                                     C.c_int32, C.c_uint32]
        lib.orc_cle_flip.restype = C.c_int
        lib.orc_regulation_matrix.argtypes = [C.POINTER(OrcNetwork), C.POINTER(OrcKinetics), _dp, _dp]
        lib.orc_regulation_loop.argtypes = [C.POINTER(OrcNetwork), C.POINTER(OrcKinetics), _dp, _dp]
        lib.orc_cle_simulate.argtypes = [C.POINTER(OrcNetwork), C.POINTER(OrcKinetics), C.POINTER(OrcRun),
                                         _dp, _dp, _dp, _dp, _up]
        lib.orc_cle_simulate.restype = C.c_int
        lib.orc_cle_population.argtypes = [C.POINTER(OrcNetwork), C.POINTER(OrcKinetics), C.POINTER(OrcRun),
                                           C.POINTER(OrcPopulation), _dp, _dp, _dp, _up]
        lib.orc_cle_population.restype = C.c_int
        lib.orc_cle_population_ex.argtypes = [C.POINTER(OrcNetwork), C.POINTER(OrcKinetics), C.POINTER(OrcRun),
                                              C.POINTER(OrcPopulation), C.c_double, _dp, _dp, _dp, _up, _up]
        lib.orc_cle_population_ex.restype = C.c_int
        _lp = C.POINTER(C.c_int64)
        lib.orc_divide.argtypes = [C.POINTER(OrcPopulation), C.c_int32, C.c_int64, C.c_uint64, C.c_uint32,
                                   C.c_int64, _dp, _dp, _dp, _lp, _lp, _dp, _dp, _dp, _dp, _lp]
        lib.orc_divide.restype = C.c_int
        lib.orc_ref_hill.argtypes = [_dp, C.c_int64, C.c_double, C.c_double, _dp, _dp]
        lib.orc_ref_init.argtypes = [C.POINTER(OrcNetwork), C.POINTER(OrcKinetics), C.POINTER(OrcRun), _dp, _dp, _up]
        lib.orc_ref_update.argtypes = [C.POINTER(OrcNetwork), C.POINTER(OrcKinetics), C.POINTER(OrcRun), C.c_int64,
                                       _dp, _dp, _dp, _dp, _up, _dp, _dp]
        lib.orc_tauleap_simulate.argtypes = lib.orc_cle_simulate.argtypes
        lib.orc_tauleap_simulate.restype = C.c_int
        lib.orc_tauleap_population.argtypes = lib.orc_cle_population.argtypes
        lib.orc_tauleap_population.restype = C.c_int
        lib.orc_poisson.argtypes = [C.c_double, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64]
        lib.orc_poisson.restype = C.c_int64
        lib.orc_binomial.argtypes = [C.c_int64, C.c_double, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64]
        lib.orc_binomial.restype = C.c_int64
        lib.orc_capture.argtypes = [_dp, C.c_int64, C.c_int32, C.c_int64, C.c_double, C.c_double, C.c_uint64,
                                    C.c_uint32, _ip, _dp]
        lib.orc_capture.restype = C.c_int
        lib.orc_moments.argtypes = [_dp, _dp, C.c_int64, C.c_int32, C.c_int64, _dp]
        _lib = lib
    return _lib


def _net(net):
    """Build an OrcNetwork from any object with the GeneNetwork attributes."""
    G = int(net.G)
    basals = np.ascontiguousarray(net.basals, dtype=np.float64)
    A = np.asfortranarray(np.asarray(net.interactions, dtype=np.float64)).ravel(order="F")
    keep = [basals, A]

    def gates(edges):
        edges = list(edges)
        if not edges:
            return 0, None, None, None, None
        tgt = np.array([e.target for e in edges], dtype=np.int32)
        off = np.zeros(len(edges) + 1, dtype=np.int32)
        off[1:] = np.cumsum([len(e.sources) for e in edges])
        src = np.array([s for e in edges for s in e.sources], dtype=np.int32)
        st = np.array([e.strength for e in edges], dtype=np.float64)
        keep.extend([tgt, off, src, st])
        return (len(edges), tgt.ctypes.data_as(_ip), off.ctypes.data_as(_ip), src.ctypes.data_as(_ip),
                st.ctypes.data_as(_dp))

    nc = gates(getattr(net, "cooperative", []))
    nr = gates(getattr(net, "redundant", []))
    kon = getattr(net, "k_on", None)
    koff = getattr(net, "k_off", None)
    konp = koffp = None
    if kon is not None:
        kon = np.ascontiguousarray(kon, dtype=np.float64)
        koff = np.ascontiguousarray(koff, dtype=np.float64)
        keep.extend([kon, koff])
        konp, koffp = kon.ctypes.data_as(_dp), koff.ctypes.data_as(_dp)
    s = OrcNetwork(G, basals.ctypes.data_as(_dp), A.ctypes.data_as(_dp), *nc, *nr, konp, koffp)
    return s, keep


def _kin(kin):
    return OrcKinetics(kin.k_t, kin.K_d, kin.n, kin.mu_m, kin.mu_p, kin.dilution)


def philox(ctr, key):
    lib = load()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.orc_philox4x32_10(c, k, o)
    return list(o)


def cle_flip(a, b, thr, seed, cell, step, g, c3base=0):
    """CLE promoter switch decision (1 = flip) of gene g from its normal words a, b."""
    return int(load().orc_cle_flip(a, b, thr, seed, cell, step, g, c3base))


def normal_pair(a, b):
    lib = load()
    z0, z1 = C.c_double(), C.c_double()
    lib.orc_normal_pair(a, b, C.byref(z0), C.byref(z1))
    return z0.value, z1.value


def regulation(net, kin, p, form="matrix"):
    lib = load()
    s, keep = _net(net)
    p = np.ascontiguousarray(p, dtype=np.float64)
    out = np.empty(net.G)
    fn = lib.orc_regulation_matrix if form == "matrix" else lib.orc_regulation_loop
    fn(C.byref(s), C.byref(_kin(kin)), p.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return out


def tauleap(net, kin, *, dt, T=None, n_steps=None, cell_num, cell_lo=0, cell_hi=None, seed=0, net_id=0,
            init="steady", snapshot_every=0, threads=1):
    """FP64 binomial tau-leap of cells [cell_lo, cell_hi) (tauleap_binomial.jl:31-180).
    Same return layout as cle(); counts as float64."""
    return cle(net, kin, dt=dt, T=T, n_steps=n_steps, cell_num=cell_num, cell_lo=cell_lo, cell_hi=cell_hi,
               seed=seed, net_id=net_id, init=init, snapshot_every=snapshot_every, threads=threads,
               _fn="orc_tauleap_simulate")


def poisson(lam, c0, c1, c3, seed):
    return int(load().orc_poisson(float(lam), c0, c1, c3, seed & 0xFFFFFFFFFFFFFFFF))


def cle(net, kin, *, dt, T=None, n_steps=None, cell_num, cell_lo=0, cell_hi=None, seed=0, net_id=0,
        init="steady", snapshot_every=0, injected=None, threads=1, _fn="orc_cle_simulate"):
    """FP64 CLE of cells [cell_lo, cell_hi).  Returns dict(m, p, snaps_m, snaps_p, promoter)
    with arrays shaped (G, n_cells) / (n_snap, G, n_cells)."""
    lib = load()
    if n_steps is None:
        n_steps = int(np.ceil(T / dt))
    cell_hi = cell_num if cell_hi is None else cell_hi
    nc = cell_hi - cell_lo
    G = net.G
    s, keep = _net(net)
    inj_p = None
    if injected is not None:
        inj = np.ascontiguousarray(injected, dtype=np.float32)
        keep.append(inj)
        inj_p = inj.ctypes.data_as(_fp)
    run = OrcRun(dt, n_steps, cell_num, cell_lo, cell_hi, seed & 0xFFFFFFFFFFFFFFFF, net_id,
                 0 if init == "steady" else 1, snapshot_every, inj_p, threads)
    n_snap = n_steps // snapshot_every if snapshot_every > 0 else 0
    m = np.empty((G, nc))
    p = np.empty((G, nc))
    sm = np.empty((max(n_snap, 1), G, nc))
    sp = np.empty((max(n_snap, 1), G, nc))
    pr = np.empty((G, nc), dtype=np.uint8)
    rc = getattr(lib, _fn)(C.byref(s), C.byref(_kin(kin)), C.byref(run), m.ctypes.data_as(_dp),
                              p.ctypes.data_as(_dp), sm.ctypes.data_as(_dp) if n_snap else None,
                              sp.ctypes.data_as(_dp) if n_snap else None, pr.ctypes.data_as(_up))
    if rc:
        raise ValueError("oracle rejected the input")
    return dict(m=m, p=p, snaps_m=sm[:n_snap], snaps_p=sp[:n_snap], promoter=pr)


def tauleap_population(net, kin, pop, *, dt, T, cell_num, seed=0, net_id=0, init="steady"):
    """FP64 binomial tau-leap with Moran population (tauleap_binomial.jl:195-261)."""
    return cle_population(net, kin, pop, dt=dt, T=T, cell_num=cell_num, seed=seed, net_id=net_id, init=init,
                          _fn="orc_tauleap_population")


def _orc_pop(pop):
    part = {"halve": 0, "binomial": 1}[pop.partition]
    rep = {"uniform": 0, "exclude_mother": 1}[pop.replace]
    return OrcPopulation(1, pop.growth_rate, pop.V_div, pop.V_init[0], pop.V_init[1], pop.div_check_interval,
                         part, rep, getattr(pop, "growth_noise", 0.0))


def divide(pop, m_gc, p_gc, V, *, seed, step, net_id=0):
    """One division check (population.jl:157-205) applied to a given state.

    m_gc, p_gc: (G, N) arrays, V: (N,).  Returns a dict with the state after
    the check (m, p, V) and the record of what it did: dividers (ascending),
    mother_m/mother_p and daughter_m/daughter_p ((n_div, G), the mothers'
    molecules after the partition and the daughters'), replaced (n_div,)."""
    lib = load()
    m = np.array(m_gc, dtype=np.float64, order="C", copy=True)
    p = np.array(p_gc, dtype=np.float64, order="C", copy=True)
    v = np.array(V, dtype=np.float64, copy=True)
    G, N = m.shape
    nd = C.c_int64()
    div = np.empty(N, dtype=np.int64)
    rep = np.empty(N, dtype=np.int64)
    mm, mp, dm, dp = (np.empty((N, G)) for _ in range(4))
    _lp = C.POINTER(C.c_int64)
    rc = lib.orc_divide(C.byref(_orc_pop(pop)), G, N, seed & 0xFFFFFFFFFFFFFFFF, net_id, int(step),
                        m.ctypes.data_as(_dp), p.ctypes.data_as(_dp), v.ctypes.data_as(_dp), C.byref(nd),
                        div.ctypes.data_as(_lp), mm.ctypes.data_as(_dp), mp.ctypes.data_as(_dp),
                        dm.ctypes.data_as(_dp), dp.ctypes.data_as(_dp), rep.ctypes.data_as(_lp))
    if rc:
        raise ValueError("oracle rejected the input")
    k = nd.value
    return dict(m=m, p=p, V=v, dividers=div[:k], mother_m=mm[:k], mother_p=mp[:k], daughter_m=dm[:k],
                daughter_p=dp[:k], replaced=rep[:k])


def cle_population(net, kin, pop, *, dt, T, cell_num, seed=0, net_id=0, init="steady", injected=None,
                   ambiguity=None, _fn="orc_cle_population"):
    """FP64 population CLE (cle.jl:143-284 with population.jl:157-205).  With
    ``ambiguity`` (a relative margin) the result also holds ``taint``: cells
    whose binomial partition rounded a value within that margin of a
    half-integer, or that hold a daughter of such a cell."""
    lib = load()
    n_steps = int(np.ceil(T / dt))
    s, keep = _net(net)
    inj_p = None
    if injected is not None:
        inj = np.ascontiguousarray(injected, dtype=np.float32)
        keep.append(inj)
        inj_p = inj.ctypes.data_as(_fp)
    run = OrcRun(dt, n_steps, cell_num, 0, cell_num, seed & 0xFFFFFFFFFFFFFFFF, net_id,
                 0 if init == "steady" else 1, 0, inj_p, 1)
    op = _orc_pop(pop)
    G = net.G
    m = np.empty((G, cell_num))
    p = np.empty((G, cell_num))
    V = np.empty(cell_num)
    pr = np.empty((G, cell_num), dtype=np.uint8)
    if ambiguity is not None:
        if _fn != "orc_cle_population":
            raise ValueError("the ambiguity margin applies to the population CLE")
        taint = np.zeros(cell_num, dtype=np.uint8)
        rc = lib.orc_cle_population_ex(C.byref(s), C.byref(_kin(kin)), C.byref(run), C.byref(op),
                                       float(ambiguity), m.ctypes.data_as(_dp), p.ctypes.data_as(_dp),
                                       V.ctypes.data_as(_dp), pr.ctypes.data_as(_up), taint.ctypes.data_as(_up))
        if rc:
            raise ValueError("oracle rejected the input")
        return dict(m=m, p=p, V=V, promoter=pr, taint=taint.astype(bool))
    rc = getattr(lib, _fn)(C.byref(s), C.byref(_kin(kin)), C.byref(run), C.byref(op),
                           m.ctypes.data_as(_dp), p.ctypes.data_as(_dp), V.ctypes.data_as(_dp),
                           pr.ctypes.data_as(_up))
    if rc:
        raise ValueError("oracle rejected the input")
    return dict(m=m, p=p, V=V, promoter=pr)


def binomial(n, p, c0, c1, c3, seed):
    return load().orc_binomial(int(n), float(p), c0, c1, c3, seed & 0xFFFFFFFFFFFFFFFF)


def capture(X_gc, efficiency, efficiency_std, seed, net_id=0):
    """X_gc: (G, cells) array.  Returns (Y (G, cells) int32, p_cell (cells,))."""
    lib = load()
    X = np.ascontiguousarray(X_gc, dtype=np.float64)
    G, cells = X.shape
    Y = np.empty((G, cells), dtype=np.int32)
    pc = np.empty(cells)
    rc = lib.orc_capture(X.ctypes.data_as(_dp), cells, G, cells, efficiency, efficiency_std,
                         seed & 0xFFFFFFFFFFFFFFFF, net_id, Y.ctypes.data_as(_ip), pc.ctypes.data_as(_dp))
    if rc:
        raise ValueError("oracle rejected the input")
    return Y, pc


def moments(m_gc, p_gc):
    lib = load()
    m = np.ascontiguousarray(m_gc, dtype=np.float64)
    p = np.ascontiguousarray(p_gc, dtype=np.float64)
    G, cells = m.shape
    out = np.empty((G, 5))
    lib.orc_moments(m.ctypes.data_as(_dp), p.ctypes.data_as(_dp), cells, G, cells, out.ctypes.data_as(_dp))
    return out
