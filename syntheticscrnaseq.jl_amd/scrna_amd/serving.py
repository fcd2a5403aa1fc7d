"""Device-resident serving loop (scrna_server_*): continuous batching of
simulate() requests of mixed G, cell count and horizon on one GPU."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .cle import CLE, SimResult, _seed
from .kinetics import CaptureModel, KineticParams
from .network import GeneNetwork


class Server:
    """Admits requests by the context's HBM budget and advances all admitted
    requests together in chunks of steps; results are identical to running
    each request alone with the same seed."""

    def __init__(self, max_slots: int = 64, ctx: _lib.Context | None = None):
        self.ctx = ctx or _lib.default_context()
        self.lib = _lib.load()
        h = C.c_void_p()
        _lib.check(self.lib.scrna_server_create(self.ctx.handle, max_slots, C.byref(h)), "scrna_server_create")
        self.handle = h
        self._meta = {}

    def submit(self, net: GeneNetwork, alg: CLE, kin: KineticParams, *, cell_num: int, T: float,
               readout: str = "protein", capture: CaptureModel | None = None, seed=None, net_id: int = 0,
               init: str = "steady") -> int:
        s, keep = net.to_c()
        run = _lib.ScrnaRun(float(alg.dt), float(T), int(cell_num), _lib.READOUT[readout],
                            0 if init == "steady" else 1, _seed(seed), int(net_id), 0, 0, None)
        cap = capture.to_c() if capture is not None else _lib.ScrnaCapture(0, 0.0, 0.0, 0)
        req = _lib.ScrnaRequest(s, kin.to_c(), run, cap)
        t = C.c_int64()
        _lib.check(self.lib.scrna_submit(self.handle, C.byref(req), C.byref(t)), "scrna_submit")
        del keep
        self._meta[t.value] = (net.G, int(cell_num), readout, capture is not None)
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1) -> SimResult:
        G, cells, readout, cap = self._meta[ticket]
        ncol = 2 * G if readout == "both" else G
        y = np.empty((cells, ncol), dtype=np.float64, order="F")
        counts = np.empty((cells, G), dtype=np.int32, order="F") if cap else None
        mom = np.empty((1, G, 5), dtype=np.float64)
        out = _lib.ScrnaOutput(y.ctypes.data_as(_lib.c_double_p), cells,
                               counts.ctypes.data_as(_lib.c_int32_p) if cap else None, None,
                               mom.ctypes.data_as(_lib.c_double_p), None)
        rc = self.lib.scrna_wait(self.handle, ticket, int(timeout_ms), C.byref(out))
        if rc != _lib.SCRNA_ERR_TIMEOUT:
            self._meta.pop(ticket, None)
        _lib.check(rc, "scrna_wait")
        return SimResult(y, counts, None, mom, None)

    def cancel(self, ticket: int) -> None:
        _lib.check(self.lib.scrna_cancel(self.handle, ticket), "scrna_cancel")

    def stats(self) -> dict:
        s = _lib.ScrnaLatency()
        _lib.check(self.lib.scrna_server_stats(self.handle, C.byref(s)), "scrna_server_stats")
        return {f: getattr(s, f) for f, _ in _lib.ScrnaLatency._fields_}

    def close(self) -> None:
        if self.handle:
            self.lib.scrna_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
