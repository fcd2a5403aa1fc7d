"""CPU baseline restatement of the reference CPU CLE -- TEST AND BASELINE
INFRASTRUCTURE ONLY.

Only tests/ and bench.py's cpu_baseline leg may import this module.  It is
timed as the CPU baseline; it is never the thing measured on the GPU and the
product never calls it.

``cle_ref`` restates src/algorithms/cle.jl:34-130 in its own structure: the
state is two G x cells matrices (Julia's column-major layout), every step
evaluates the Hill terms of all cells in one fused pass (cle.jl:79-82),
forms the regulation as two dense products A_pos * act + A_neg * rep
(cle.jl:84-86; numpy's matmul calls OpenBLAS dgemm, as Julia's mul! does,
on the BLAS threads), then one fused pass adds the AND/OR corrections
(cle.jl:89-106) and applies the noise and the mRNA and protein updates
(cle.jl:108-120, protein from the updated mRNA), like Julia's fused @.
broadcasts on one thread.  The fused passes are C loops in liboracle.so
(orc_ref_hill, orc_ref_update) that draw the oracle's Philox stream, so a
cle_ref run reproduces orc_cle_simulate's (tests/test_cpu_ref.py).
Telegraph promoters (no reference CLE has them) follow the oracle's rules.
"""
from __future__ import annotations

import numpy as np


def _split(net):
    """A_pos, A_neg from A_eff = A .* beta (rows), zero diagonal (network.jl:206-219)."""
    A_eff = np.asarray(net.interactions, dtype=np.float64) * np.asarray(net.basals, dtype=np.float64)[:, None]
    np.fill_diagonal(A_eff, 0.0)
    return np.ascontiguousarray(np.maximum(A_eff, 0.0)), np.ascontiguousarray(np.maximum(-A_eff, 0.0))


def cle_ref(net, kin, *, dt, n_steps, cells, seed=0, net_id=0, init="steady", injected=None, threads=1,
            noise="philox"):
    """Run the cell-vectorised CLE on `cells` cells; returns (M, P), each G x cells.

    noise="philox" draws the oracle's stream (reproducible against it);
    noise="randn" fills two noise matrices per step from numpy's ziggurat
    normals, as cle.jl:109, 116 do with randn! -- the cost model of the
    reference, used for timing (telegraph switching still reads Philox)."""
    import ctypes as C

    from threadpoolctl import threadpool_limits

    from . import oracle as orc

    lib = orc.load()
    G = int(net.G)
    A_pos, A_neg = _split(net)
    A_posT = np.ascontiguousarray(A_pos.T)
    A_negT = np.ascontiguousarray(A_neg.T)
    s_net, keep = orc._net(net)
    s_kin = orc._kin(kin)
    inj_p = None
    if injected is not None:
        inj = np.ascontiguousarray(injected, dtype=np.float32)
        keep.append(inj)
        inj_p = inj.ctypes.data_as(orc._fp)
    run = orc.OrcRun(dt, n_steps, cells, 0, cells, seed & 0xFFFFFFFFFFFFFFFF, net_id, 0 if init == "steady" else 1,
                     0, inj_p, 1)
    # (cells, G) row-major arrays are Julia's G x cells column-major matrices
    M = np.empty((cells, G))
    P = np.empty((cells, G))
    act = np.empty((cells, G))
    rep = np.empty((cells, G))
    reg = np.empty((cells, G))
    tmp = np.empty((cells, G))
    on = np.empty((cells, G), dtype=np.uint8)
    dp = lambda a: a.ctypes.data_as(orc._dp)
    lib.orc_ref_init(C.byref(s_net), C.byref(s_kin), C.byref(run), dp(M), dp(P), on.ctypes.data_as(orc._up))
    Kn = kin.K_d ** kin.n
    nm = npn = None
    nullp = None
    if noise == "randn":
        gen = np.random.default_rng(seed)
        nm = np.empty((cells, G))
        npn = np.empty((cells, G))
    elif noise != "philox":
        raise ValueError("noise must be 'philox' or 'randn'")
    with threadpool_limits(limits=threads):
        for s in range(n_steps):
            lib.orc_ref_hill(dp(P), cells * G, kin.n, Kn, dp(act), dp(rep))     # cle.jl:79-82
            np.matmul(act, A_posT, out=reg)                                     # cle.jl:84 (act' A_pos' = (A_pos act)')
            np.matmul(rep, A_negT, out=tmp)                                     # cle.jl:85
            reg += tmp                                                          # cle.jl:86
            if nm is not None:
                gen.standard_normal(out=nm)                                     # cle.jl:109
                gen.standard_normal(out=npn)                                    # cle.jl:116
            lib.orc_ref_update(C.byref(s_net), C.byref(s_kin), C.byref(run), s, dp(M), dp(P), dp(reg), dp(act),
                               on.ctypes.data_as(orc._up), dp(nm) if nm is not None else nullp,
                               dp(npn) if npn is not None else nullp)           # cle.jl:89-120
    del keep
    return M.T.copy(), P.T.copy()
