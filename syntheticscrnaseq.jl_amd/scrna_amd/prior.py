"""Network prior sampler and theta constructor for simulation-based inference.

``sample_network(G, seed=..., index=...)`` replaces the reference's
``sample_network(G; rng, regulation, ...)`` (src/kinetics.jl:105-181) and
``network_from_theta(theta, G)`` replaces ``GeneNetwork(theta, G)``
(src/network.jl:233-256).  Both run in libscrna.so's native code (host side);
network ``index`` of a ``seed`` is a pure function of the two, so SBI batches
can be drawn in any order and fed straight to ``simulate_gpu_batch``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .network import CoopEdge, GeneNetwork, RedunEdge

REGULATION = {"mixed": 0, "additive": 1, "cooperative": 2, "redundant": 3}


def _range(r):
    if r is None:
        return -1, -1
    lo, hi = (r.start, r.stop - 1) if isinstance(r, range) else (int(r[0]), int(r[1]))
    if lo < 0 or hi < lo:
        raise ValueError(f"invalid range {r}")
    return lo, hi


def _prior(G, regulation, n_tf_range, targets_per_tf_range, n_coop_range, n_redun_range, infer_basals):
    if regulation not in REGULATION:
        raise ValueError(f"regulation must be one of {sorted(REGULATION)}")
    return _lib.ScrnaPrior(int(G), REGULATION[regulation], *_range(n_tf_range), *_range(targets_per_tf_range),
                           *_range(n_coop_range), *_range(n_redun_range), 1 if infer_basals else 0)


def sample_network(G: int, *, seed: int, index: int = 0, regulation: str = "mixed", n_tf_range=None,
                   targets_per_tf_range=None, n_coop_range=None, n_redun_range=None,
                   infer_basals: bool = True) -> GeneNetwork:
    """One network from the prior.  Ranges are inclusive ``(lo, hi)`` tuples or
    Python ranges (``range(1, 3)`` is Julia's ``1:2``); ``None`` takes the
    reference's G-dependent default (kinetics.jl:107-111)."""
    lib = _lib.load()
    pr = _prior(G, regulation, n_tf_range, targets_per_tf_range, n_coop_range, n_redun_range, infer_basals)
    mg = C.c_int32(0)
    _lib.check(lib.scrna_prior_max_gates(C.byref(pr), C.byref(mg)), "prior")
    m = mg.value
    basals = np.empty(G)
    A = np.empty((G, G), order="F")
    gtype = np.empty(m, dtype=np.int32)
    gtarget = np.empty(m, dtype=np.int32)
    gsrc = np.empty(2 * m, dtype=np.int32)
    gstr = np.empty(m)
    ng = C.c_int32(0)
    rc = lib.scrna_sample_network(C.byref(pr), int(seed) & 0xFFFFFFFFFFFFFFFF, int(index),
                                  basals.ctypes.data_as(_lib.c_double_p), A.ctypes.data_as(_lib.c_double_p),
                                  gtype.ctypes.data_as(_lib.c_int32_p), gtarget.ctypes.data_as(_lib.c_int32_p),
                                  gsrc.ctypes.data_as(_lib.c_int32_p), gstr.ctypes.data_as(_lib.c_double_p), m,
                                  C.byref(ng))
    _lib.check(rc, "sample_network")
    coop, redun = [], []
    for k in range(ng.value):
        e = (int(gtarget[k]), (int(gsrc[2 * k]), int(gsrc[2 * k + 1])), float(gstr[k]))
        (coop if gtype[k] == 0 else redun).append(CoopEdge(*e) if gtype[k] == 0 else RedunEdge(*e))
    return GeneNetwork(G, basals, np.ascontiguousarray(A), coop, redun)


def sample_networks(G: int, n: int, *, seed: int, first_index: int = 0, **kw) -> list:
    """Networks ``first_index .. first_index + n - 1`` of ``seed``."""
    return [sample_network(G, seed=seed, index=first_index + i, **kw) for i in range(n)]


def network_from_theta(theta, G: int) -> GeneNetwork:
    """GeneNetwork(theta, G) (network.jl:233-256)."""
    lib = _lib.load()
    th = np.ascontiguousarray(theta, dtype=np.float64)
    basals = np.empty(G)
    A = np.empty((G, G), order="F")
    rc = lib.scrna_network_from_theta(int(G), th.ctypes.data_as(_lib.c_double_p), th.size,
                                      basals.ctypes.data_as(_lib.c_double_p), A.ctypes.data_as(_lib.c_double_p))
    _lib.check(rc, "network_from_theta")
    return GeneNetwork(G, basals, np.ascontiguousarray(A))
