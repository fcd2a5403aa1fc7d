"""Gene regulatory network types (reference: src/network.jl).

Mirrors ``CoopEdge``, ``RedunEdge`` and ``GeneNetwork`` with the same field
names and meaning.  Indices are 0-based (Python), where the Julia reference
is 1-based.  ``interactions[i, j]`` is the effect of protein j on gene i
(network.jl:43-45); the diagonal must be zero (network.jl:68).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from . import _lib


@dataclass(frozen=True)
class CoopEdge:
    """AND gate: ``strength * beta[target] * prod(hill_act(p[src]))`` (network.jl:16-20)."""
    target: int
    sources: tuple
    strength: float

    def __init__(self, target: int, sources: Sequence[int], strength: float):
        object.__setattr__(self, "target", int(target))
        object.__setattr__(self, "sources", tuple(int(s) for s in sources))
        object.__setattr__(self, "strength", float(strength))


@dataclass(frozen=True)
class RedunEdge:
    """OR gate: ``strength * beta[target] * (1 - prod(1 - hill_act(p[src])))`` (network.jl:30-34)."""
    target: int
    sources: tuple
    strength: float

    def __init__(self, target: int, sources: Sequence[int], strength: float):
        object.__setattr__(self, "target", int(target))
        object.__setattr__(self, "sources", tuple(int(s) for s in sources))
        object.__setattr__(self, "strength", float(strength))


@dataclass
class GeneNetwork:
    """GeneNetwork(G, basals, interactions[, cooperative, redundant, k_on, k_off]) (network.jl:55-96).

    ``k_on`` / ``k_off`` default to +inf (constitutive, network.jl:48-49).
    Construction asserts the same invariants as the reference constructors.
    """
    G: int
    basals: np.ndarray
    interactions: np.ndarray
    cooperative: list = field(default_factory=list)
    redundant: list = field(default_factory=list)
    k_on: np.ndarray | None = None
    k_off: np.ndarray | None = None

    def __post_init__(self):
        self.G = int(self.G)
        self.basals = np.ascontiguousarray(self.basals, dtype=np.float64)
        self.interactions = np.asarray(self.interactions, dtype=np.float64)
        if self.basals.shape != (self.G,):                        # network.jl:66
            raise ValueError(f"basals must have length G={self.G}")
        if self.interactions.shape != (self.G, self.G):           # network.jl:67
            raise ValueError(f"interactions must be {self.G}x{self.G}")
        if np.any(np.diag(self.interactions) != 0.0):             # network.jl:68
            raise ValueError("interaction diagonal must be zero")
        self.k_on = (np.full(self.G, np.inf) if self.k_on is None
                     else np.ascontiguousarray(self.k_on, dtype=np.float64))
        self.k_off = (np.full(self.G, np.inf) if self.k_off is None
                      else np.ascontiguousarray(self.k_off, dtype=np.float64))
        if self.k_on.shape != (self.G,) or self.k_off.shape != (self.G,):   # network.jl:93-94
            raise ValueError("k_on and k_off must have length G")
        for e in list(self.cooperative) + list(self.redundant):
            if not (0 <= e.target < self.G) or any(not (0 <= s < self.G) for s in e.sources):
                raise ValueError(f"gate {e} refers to a gene outside 0..{self.G - 1}")
        self._c_cache = None

    @classmethod
    def from_basals(cls, basals, interactions, *, cooperative=(), redundant=(), k_on=None, k_off=None):
        """Keyword constructor that infers G from basals (network.jl:85-96)."""
        basals = np.asarray(basals, dtype=np.float64)
        return cls(len(basals), basals, interactions, list(cooperative), list(redundant), k_on, k_off)

    # -- ctypes view --------------------------------------------------------
    def to_c(self) -> tuple:
        """Return (ScrnaNetwork, keepalive) for the C ABI (column-major A)."""
        A = np.asfortranarray(self.interactions).ravel(order="F")
        keep = [self.basals, A]

        def gates(edges):
            if not edges:
                return 0, None, None, None, None
            tgt = np.array([e.target for e in edges], dtype=np.int32)
            off = np.zeros(len(edges) + 1, dtype=np.int32)
            off[1:] = np.cumsum([len(e.sources) for e in edges])
            src = np.array([s for e in edges for s in e.sources], dtype=np.int32)
            st = np.array([e.strength for e in edges], dtype=np.float64)
            keep.extend([tgt, off, src, st])
            return (len(edges), tgt.ctypes.data_as(_lib.c_int32_p), off.ctypes.data_as(_lib.c_int32_p),
                    src.ctypes.data_as(_lib.c_int32_p), st.ctypes.data_as(_lib.c_double_p))

        nc, ct, co, cs, cst = gates(self.cooperative)
        nr, rt, ro, rs, rst = gates(self.redundant)
        kon = np.ascontiguousarray(self.k_on, dtype=np.float64)
        koff = np.ascontiguousarray(self.k_off, dtype=np.float64)
        keep.extend([kon, koff])
        s = _lib.ScrnaNetwork(
            self.G, self.basals.ctypes.data_as(_lib.c_double_p), A.ctypes.data_as(_lib.c_double_p),
            nc, ct, co, cs, cst, nr, rt, ro, rs, rst,
            kon.ctypes.data_as(_lib.c_double_p), koff.ctypes.data_as(_lib.c_double_p))
        return s, keep


def is_bursty(net: GeneNetwork) -> bool:
    """True if any gene has a finite switching rate (network.jl:103)."""
    return bool(np.any(np.isfinite(net.k_on)) or np.any(np.isfinite(net.k_off)))


def n_species(net: GeneNetwork) -> int:
    """2G: G mRNAs plus G proteins (network.jl:263)."""
    return 2 * net.G


def hill_activation(p, K, n):
    """p^n / (p^n + K^n) (network.jl:112-115)."""
    pn = np.power(p, n)
    return pn / (pn + K ** n)


def hill_repression(p, K, n):
    """K^n / (p^n + K^n) (network.jl:122-125)."""
    pn = np.power(p, n)
    return K ** n / (pn + K ** n)


def precompute_hill_matrices(net: GeneNetwork):
    """(A_pos, A_neg) with A_eff[i,j] = A[i,j]*beta[i], zero diagonal (network.jl:206-219)."""
    A_eff = net.interactions * net.basals[:, None]
    np.fill_diagonal(A_eff, 0.0)
    return np.maximum(A_eff, 0.0), np.maximum(-A_eff, 0.0)


def gene_network_from_theta(theta, G: int) -> GeneNetwork:
    """GeneNetwork(theta, G): log-basals then row-major off-diagonal A (network.jl:233-256)."""
    theta = np.asarray(theta, dtype=np.float64)
    n_int = G * (G - 1)
    if theta.size == n_int + G:
        basals, a = np.exp(theta[:G]), theta[G:]
    elif theta.size == n_int:
        basals, a = np.full(G, np.exp(-1.0)), theta
    else:
        raise ValueError(f"theta must have length {n_int} or {n_int + G}, got {theta.size}")
    A = np.zeros((G, G))
    A[~np.eye(G, dtype=bool)] = a
    return GeneNetwork(G, basals, A)
