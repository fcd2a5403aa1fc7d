"""ctypes binding of libscrna.so (the C ABI declared in include/scrna.h).

The shared library is the product: every simulation in this package runs in
its HIP kernels.  There is no CPU fallback -- if the library or a gfx950
device is missing, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
# SCRNA_LIB overrides the in-tree library (as in the Julia shim), e.g. to compare builds.
LIB_PATH = os.environ.get("SCRNA_LIB") or os.path.join(os.path.dirname(_HERE), "libscrna.so")

SCRNA_OK = 0
SCRNA_ERR_INVALID = -1
SCRNA_ERR_BUDGET = -2
SCRNA_ERR_DEVICE = -3
SCRNA_ERR_CANCELLED = -4
SCRNA_ERR_TIMEOUT = -5
SCRNA_ERR_UNSUPPORTED = -6

READOUT = {"mrna": 0, "protein": 1, "both": 2}
PRECISION = {"fp32": 0, "bf16": 1, "bf16x3": 2, "fp32_unfused": 3, "fp32_mfma": 4, "fp32_valu": 5}

c_double_p = C.POINTER(C.c_double)
c_int32_p = C.POINTER(C.c_int32)
c_float_p = C.POINTER(C.c_float)


class ScrnaNetwork(C.Structure):
    _fields_ = [
        ("G", C.c_int32),
        ("basals", c_double_p),
        ("A", c_double_p),
        ("n_coop", C.c_int32),
        ("coop_target", c_int32_p),
        ("coop_off", c_int32_p),
        ("coop_src", c_int32_p),
        ("coop_strength", c_double_p),
        ("n_redun", C.c_int32),
        ("redun_target", c_int32_p),
        ("redun_off", c_int32_p),
        ("redun_src", c_int32_p),
        ("redun_strength", c_double_p),
        ("k_on", c_double_p),
        ("k_off", c_double_p),
    ]


class ScrnaKinetics(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("k_t", "K_d", "n", "mu_m", "mu_p", "dilution")]


class ScrnaPopulation(C.Structure):
    _fields_ = [
        ("enabled", C.c_int32),
        ("growth_rate", C.c_double),
        ("V_div", C.c_double),
        ("V_init_lo", C.c_double),
        ("V_init_hi", C.c_double),
        ("div_check_interval", C.c_int32),
        ("partition", C.c_int32),
        ("replace_mode", C.c_int32),
        ("growth_noise", C.c_double),
    ]


class ScrnaCapture(C.Structure):
    _fields_ = [
        ("enabled", C.c_int32),
        ("efficiency", C.c_double),
        ("efficiency_std", C.c_double),
        ("readout", C.c_int32),
    ]


class ScrnaRun(C.Structure):
    _fields_ = [
        ("dt", C.c_double),
        ("T", C.c_double),
        ("cell_num", C.c_int64),
        ("readout", C.c_int32),
        ("init_mode", C.c_int32),
        ("seed", C.c_uint64),
        ("net_id", C.c_uint32),
        ("precision", C.c_int32),
        ("snapshot_every", C.c_int64),
        ("injected_noise", c_float_p),
        ("algorithm", C.c_int32),
        ("reserved0", C.c_int32),
    ]


class ScrnaOutput(C.Structure):
    _fields_ = [
        ("y", c_double_p),
        ("ld", C.c_int64),
        ("counts", c_int32_p),
        ("snapshots", c_float_p),
        ("moments", c_double_p),
        ("volumes", c_double_p),
    ]


class ScrnaRequest(C.Structure):
    _fields_ = [
        ("net", ScrnaNetwork),
        ("kin", ScrnaKinetics),
        ("run", ScrnaRun),
        ("capture", ScrnaCapture),
    ]


class ScrnaPrior(C.Structure):
    _fields_ = [
        ("G", C.c_int32),
        ("regulation", C.c_int32),
        ("n_tf_lo", C.c_int32),
        ("n_tf_hi", C.c_int32),
        ("targets_lo", C.c_int32),
        ("targets_hi", C.c_int32),
        ("coop_lo", C.c_int32),
        ("coop_hi", C.c_int32),
        ("redun_lo", C.c_int32),
        ("redun_hi", C.c_int32),
        ("infer_basals", C.c_int32),
    ]


class ScrnaLatency(C.Structure):
    _fields_ = [
        ("p50_ms", C.c_double),
        ("p99_ms", C.c_double),
        ("mean_ms", C.c_double),
        ("max_ms", C.c_double),
        ("cell_gene_steps", C.c_double),
        ("busy_s", C.c_double),
        ("completed", C.c_int64),
        ("rejected", C.c_int64),
        ("cancelled", C.c_int64),
    ]


# Every entry point of include/scrna.h with its ctypes signature.
_VP = C.c_void_p
SIGNATURES = {
    "scrna_abi_version": (C.c_int32, []),
    "scrna_version": (C.c_char_p, []),
    "scrna_last_error": (C.c_char_p, []),
    "scrna_ctx_create": (C.c_int, [C.c_int32, C.c_uint64, C.POINTER(_VP)]),
    "scrna_ctx_destroy": (C.c_int, [_VP]),
    "scrna_simulate_cle": (C.c_int, [_VP, C.POINTER(ScrnaNetwork), C.POINTER(ScrnaKinetics),
                                     C.POINTER(ScrnaRun), C.POINTER(ScrnaPopulation),
                                     C.POINTER(ScrnaCapture), C.POINTER(ScrnaOutput)]),
    "scrna_simulate_cle_noalloc": (C.c_int, [_VP, C.POINTER(ScrnaNetwork), C.POINTER(ScrnaKinetics),
                                             C.POINTER(ScrnaRun), C.POINTER(ScrnaCapture), C.POINTER(ScrnaOutput)]),
    "scrna_noalloc_scratch_bytes": (C.c_uint64, []),
    "scrna_simulate_cle_batch": (C.c_int, [_VP, C.c_int32, C.POINTER(ScrnaNetwork),
                                           C.POINTER(ScrnaKinetics), C.POINTER(ScrnaRun),
                                           C.POINTER(ScrnaPopulation), C.POINTER(ScrnaCapture),
                                           C.POINTER(ScrnaOutput)]),
    "scrna_apply_capture": (C.c_int, [_VP, c_double_p, C.c_int64, C.c_int32, C.c_int64, C.c_double,
                                      C.c_double, C.c_uint64, C.c_uint32, c_int32_p]),
    "scrna_plan_create": (C.c_int, [_VP, C.c_int32, C.POINTER(ScrnaNetwork), C.POINTER(ScrnaKinetics),
                                    C.POINTER(ScrnaRun), C.POINTER(ScrnaPopulation),
                                    C.POINTER(ScrnaCapture), C.POINTER(_VP)]),
    "scrna_plan_launch": (C.c_int, [_VP, _VP]),
    "scrna_plan_sync": (C.c_int, [_VP]),
    "scrna_plan_download": (C.c_int, [_VP, C.c_int32, C.POINTER(ScrnaOutput)]),
    "scrna_plan_stepper_ms": (C.c_int, [_VP, c_double_p, c_double_p]),
    "scrna_plan_bytes": (C.c_int, [_VP, C.POINTER(C.c_uint64)]),
    "scrna_plan_use_graph": (C.c_int, [_VP]),
    "scrna_plan_destroy": (C.c_int, [_VP]),
    "scrna_autotune": (C.c_int, [_VP, C.c_int32, C.c_char_p, C.POINTER(C.c_int32)]),
    "scrna_autotune_phases": (C.c_int, [_VP, C.c_int32, C.c_int32, C.c_char_p, C.POINTER(C.c_int32)]),
    "scrna_ctx_set_tile": (C.c_int, [_VP, C.c_int32, C.c_int32, C.c_int32]),
    "scrna_server_create": (C.c_int, [_VP, C.c_int32, C.POINTER(_VP)]),
    "scrna_submit": (C.c_int, [_VP, C.POINTER(ScrnaRequest), C.POINTER(C.c_int64)]),
    "scrna_wait": (C.c_int, [_VP, C.c_int64, C.c_int64, C.POINTER(ScrnaOutput)]),
    "scrna_cancel": (C.c_int, [_VP, C.c_int64]),
    "scrna_server_stats": (C.c_int, [_VP, C.POINTER(ScrnaLatency)]),
    "scrna_server_destroy": (C.c_int, [_VP]),
    "scrna_prior_max_gates": (C.c_int, [C.POINTER(ScrnaPrior), c_int32_p]),
    "scrna_sample_network": (C.c_int, [C.POINTER(ScrnaPrior), C.c_uint64, C.c_uint32, c_double_p, c_double_p,
                                       c_int32_p, c_int32_p, c_int32_p, c_double_p, C.c_int32, c_int32_p]),
    "scrna_network_from_theta": (C.c_int, [C.c_int32, c_double_p, C.c_int64, c_double_p, c_double_p]),
    "scrna_write_10x": (C.c_int, [C.c_char_p, c_int32_p, C.c_int64, C.c_int32, C.c_int64,
                                  C.POINTER(C.c_char_p), C.c_char_p, C.POINTER(C.c_int64)]),
}

_lib = None
_lock = threading.Lock()


class ScrnaError(RuntimeError):
    """Device or runtime failure reported by libscrna."""


class BudgetError(MemoryError):
    """The request does not fit the context's HBM budget."""


class UnsupportedError(NotImplementedError):
    """The combination is not supported (the reference raises error())."""


def load(path: str | None = None):
    """Load libscrna.so; raise if it has not been built."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise ScrnaError(f"libscrna.so not built at {p}; run __graft_entry__.build() or "
                             f"python syntheticscrnaseq.jl_amd/build.py")
        lib = C.CDLL(p)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if path is None:
            _lib = lib
        return lib


def check(rc: int, what: str = "") -> None:
    if rc == SCRNA_OK:
        return
    msg = (load().scrna_last_error() or b"").decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == SCRNA_ERR_INVALID:
        raise ValueError(text)
    if rc == SCRNA_ERR_BUDGET:
        raise BudgetError(text)
    if rc == SCRNA_ERR_UNSUPPORTED:
        raise UnsupportedError(text)
    if rc == SCRNA_ERR_TIMEOUT:
        raise TimeoutError(text)
    raise ScrnaError(f"[{rc}] {text}")


class Context:
    """One device and one HBM budget (scrna_ctx)."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.scrna_ctx_create(device, hbm_budget_bytes, C.byref(h)), "scrna_ctx_create")
        self.handle = h

    def autotune(self, max_G: int = 256, cache_path: str | None = None, phase_steps: int = 0) -> int:
        """Pick genes-per-lane per G bucket by timing (scrna_autotune); returns buckets tuned.
        phase_steps > 0 times shapes as population plans with that division interval run them
        (scrna_autotune_phases)."""
        n = C.c_int32()
        path = cache_path.encode() if cache_path else None
        check(self.lib.scrna_autotune_phases(self.handle, int(max_G), int(phase_steps), path, C.byref(n)),
              "scrna_autotune_phases")
        return n.value

    def set_tile(self, G_lo: int, G_hi: int, genes_per_lane: int) -> None:
        check(self.lib.scrna_ctx_set_tile(self.handle, G_lo, G_hi, genes_per_lane), "scrna_ctx_set_tile")

    def close(self) -> None:
        if self.handle:
            self.lib.scrna_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Context | None = None


def default_context() -> Context:
    """The process-wide context on device 0 (created on first use)."""
    global _default_ctx
    if _default_ctx is None:
        ctx = Context(0)
        with _lock:
            if _default_ctx is None:
                _default_ctx = ctx
    return _default_ctx
