"""Plain-Python restatement of the network prior sampler -- TEST
INFRASTRUCTURE ONLY (the checker for libscrna's scrna_sample_network; only
tests/ may import it).

Follows sample_network (src/kinetics.jl:105-181) and _sample_pair_edge
(src/kinetics.jl:183-205) step by step, drawing from the same keyed stream as
the library: Philox4x32-10 with key = seed and counter (index, block, 0, 9),
two 52-bit uniforms per block, consumed in the order documented in prior.cpp.
GeneNetwork(theta, G) (src/network.jl:233-256) is restated too.
"""
from __future__ import annotations

import math

import numpy as np

from . import oracle as orc

PURPOSE = 9


class Stream:
    def __init__(self, seed: int, index: int):
        self.key = (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
        self.index = index
        self.block = 0
        self.buf = []
        self.spare = None

    def uniform(self) -> float:
        if not self.buf:
            u = orc.philox((self.index, self.block, 0, PURPOSE), self.key)
            self.block += 1
            self.buf = [(u[0], u[1]), (u[2], u[3])]
        a, b = self.buf.pop(0)
        x = (a << 20) | (b >> 12)
        return (x + 0.5) * 2.0 ** -52

    def range(self, lo: int, hi: int) -> int:
        return min(lo + int(math.floor(self.uniform() * (hi - lo + 1))), hi)

    def normal(self) -> float:
        if self.spare is not None:
            z, self.spare = self.spare, None
            return z
        u1, u2 = self.uniform(), self.uniform()
        r = math.sqrt(-2.0 * math.log(u1))
        self.spare = r * math.sin(6.283185307179586 * u2)
        return r * math.cos(6.283185307179586 * u2)

    def strength(self) -> float:                   # kinetics.jl:140-144
        neg = self.uniform() < 0.5
        u = self.uniform()
        return -10.0 + 7.0 * u if neg else 3.0 + 7.0 * u

    def choose(self, pool, k):                     # sampling without replacement, selection order
        pool = list(pool)
        n = len(pool)
        k = min(k, n)
        for i in range(k):
            j = i + min(int(math.floor(self.uniform() * (n - i))), n - i - 1)
            pool[i], pool[j] = pool[j], pool[i]
        return pool[:k]


def defaults(G):                                    # kinetics.jl:107-111
    return dict(n_tf=(1, 2) if G <= 5 else (2, 3) if G <= 10 else (4, 5),
                targets=(1, 3) if G <= 5 else (2, 4) if G <= 10 else (3, 6),
                coop=(1, 2) if G > 10 else (0, 1), redun=(1, 2) if G > 10 else (0, 1))


def _pair_edge(G, used, rs):                        # kinetics.jl:183-205
    if G < 3:
        return None
    for _ in range(20):
        src = sorted(rs.choose(range(G), 2))
        targets = [g for g in range(G) if g not in src]
        t = targets[rs.range(0, len(targets) - 1)]
        if (t, src[0]) not in used and (t, src[1]) not in used:
            used.add((t, src[0]))
            used.add((t, src[1]))
            return t, tuple(src), rs.strength()
    return None


def sample_network(G, seed, index=0, regulation="mixed", n_tf=None, targets=None, coop=None, redun=None,
                   infer_basals=True):
    """Returns (basals, A, gates) with gates = [(type, target, (s1, s2), strength)], type 0 AND / 1 OR."""
    d = defaults(G)
    n_tf = n_tf or d["n_tf"]
    targets = targets or d["targets"]
    coop = coop or d["coop"]
    redun = redun or d["redun"]
    rs = Stream(seed, index)
    basals = np.array([math.exp(rs.normal() * 0.5 - 1.0) if infer_basals else math.exp(-1.0)
                       for _ in range(G)])                                     # kinetics.jl:115-119
    A = np.zeros((G, G))
    used = set()
    tfs = sorted(rs.choose(range(G), min(rs.range(*n_tf), G)))                 # kinetics.jl:123-124
    for tf in tfs:                                                             # kinetics.jl:127-147
        want = rs.range(*targets)
        for t in rs.choose([g for g in range(G) if g != tf], want):
            A[t, tf] = rs.strength()
            used.add((t, tf))
    gates = []
    if regulation in ("mixed", "cooperative"):                                 # kinetics.jl:153-162
        for _ in range(rs.range(*coop)):
            e = _pair_edge(G, used, rs)
            if e:
                gates.append((0,) + e)
    if regulation in ("mixed", "redundant"):                                   # kinetics.jl:164-173
        for _ in range(rs.range(*redun)):
            e = _pair_edge(G, used, rs)
            if e:
                gates.append((1,) + e)
    if regulation == "mixed" and not gates:                                    # kinetics.jl:169-175
        e = _pair_edge(G, used, rs)
        if e:
            gates.append((1,) + e)
    return basals, A, gates


def network_from_theta(theta, G):                   # network.jl:233-256
    theta = np.asarray(theta, dtype=np.float64)
    n_int = G * (G - 1)
    if theta.size == n_int + G:
        basals, a = np.exp(theta[:G]), theta[G:]
    elif theta.size == n_int:
        basals, a = np.full(G, math.exp(-1.0)), theta
    else:
        raise ValueError("theta length")
    A = np.zeros((G, G))
    idx = 0
    for i in range(G):
        for j in range(G):
            if i != j:
                A[i, j] = a[idx]
                idx += 1
    return basals, A
