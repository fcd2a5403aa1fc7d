"""C-ABI checks that need no GPU: the library loads, exports every entry
point include/scrna.h declares, and the ctypes structs match the C layout."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from scrna_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "scrna.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(scrna_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_signatures_cover_header():
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_version_calls_without_device():
    lib = _lib.load()
    assert lib.scrna_abi_version() == 3
    assert b"gfx950" in lib.scrna_version()


STRUCTS = {
    "scrna_network": (_lib.ScrnaNetwork, ["G", "basals", "A", "n_coop", "coop_target", "coop_off", "coop_src",
                                          "coop_strength", "n_redun", "redun_target", "redun_off", "redun_src",
                                          "redun_strength", "k_on", "k_off"]),
    "scrna_kinetics": (_lib.ScrnaKinetics, ["k_t", "K_d", "n", "mu_m", "mu_p", "dilution"]),
    "scrna_population": (_lib.ScrnaPopulation, ["enabled", "growth_rate", "V_div", "V_init_lo", "V_init_hi",
                                                "div_check_interval", "partition", "replace_mode", "growth_noise"]),
    "scrna_capture": (_lib.ScrnaCapture, ["enabled", "efficiency", "efficiency_std", "readout"]),
    "scrna_run": (_lib.ScrnaRun, ["dt", "T", "cell_num", "readout", "init_mode", "seed", "net_id", "precision",
                                  "snapshot_every", "injected_noise", "algorithm", "reserved0"]),
    "scrna_output": (_lib.ScrnaOutput, ["y", "ld", "counts", "snapshots", "moments", "volumes"]),
    "scrna_request": (_lib.ScrnaRequest, ["net", "kin", "run", "capture"]),
    "scrna_latency": (_lib.ScrnaLatency, ["p50_ms", "p99_ms", "mean_ms", "max_ms", "cell_gene_steps", "busy_s",
                                          "completed", "rejected", "cancelled"]),
    "scrna_prior": (_lib.ScrnaPrior, ["G", "regulation", "n_tf_lo", "n_tf_hi", "targets_lo", "targets_hi",
                                      "coop_lo", "coop_hi", "redun_lo", "redun_hi", "infer_basals"]),
}


def test_struct_layout_matches_c():
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void){"]
    for cname, (_, fields) in STRUCTS.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f in fields:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "l.c")
        exe = os.path.join(d, "l")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-o", exe, src], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(line.split() for line in out if line)
    for cname, (cls, fields) in STRUCTS.items():
        assert int(got[cname]) == C.sizeof(cls), cname
        for f in fields:
            assert int(got[f"{cname}.{f}"]) == getattr(cls, f).offset, f"{cname}.{f}"


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(_lib.ScrnaError):
        _lib.load(str(tmp_path / "nope.so"))


def test_julia_shim_binds_only_exported_symbols():
    """Every ccall of the Julia shim names a symbol the library exports and the
    header declares (the shim cannot run here: no Julia)."""
    shim = os.path.join(ROOT, "syntheticscrnaseq.jl_amd", "julia", "SyntheticscRNAseqHIP.jl")
    called = set(re.findall(r"ccall\(\(:(\w+), LIB\)", open(shim).read()))
    assert {"scrna_simulate_cle", "scrna_simulate_cle_batch", "scrna_simulate_cle_noalloc"} <= called
    lib = C.CDLL(_lib.LIB_PATH)
    declared = set(declared_functions())
    for name in sorted(called):
        assert name in declared, name
        assert hasattr(lib, name), name
