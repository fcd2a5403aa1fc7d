"""Host-side mirror of SyntheticscRNAseq.jl's GPU CLE interface over libscrna.so.

The names follow the reference package (src/SyntheticscRNAseq.jl:40-79):
GeneNetwork, CoopEdge, RedunEdge, KineticParams, steady_state,
PopulationConfig, CaptureModel, CLE, simulate, simulate_gpu_batch,
apply_capture, and the validation helpers ks_statistic, ks_critical_value,
compare_moments, convergence_rate.  Every simulation runs in the HIP kernels of libscrna.so on a
gfx950 device; nothing here computes a trajectory on the CPU.
"""
from . import analytical, io, prior, samplers, validate
from ._lib import BudgetError, Context, ScrnaError, UnsupportedError, default_context, load
from .cle import CLE, BinomialTauLeap, Plan, SimResult, apply_capture, n_steps, simulate, simulate_gpu_batch
from .kinetics import CaptureModel, KineticParams, PopulationConfig, steady_state
from .network import (CoopEdge, GeneNetwork, RedunEdge, gene_network_from_theta, hill_activation,
                      hill_repression, is_bursty, n_species, precompute_hill_matrices)
from .serving import Server
from .validate import compare_moments, convergence_rate, ks_critical_value, ks_statistic

__all__ = [
    "BudgetError", "Context", "ScrnaError", "UnsupportedError", "default_context", "load",
    "CLE", "BinomialTauLeap", "Plan", "SimResult", "apply_capture", "n_steps", "simulate", "simulate_gpu_batch",
    "CaptureModel", "KineticParams", "PopulationConfig", "steady_state",
    "CoopEdge", "GeneNetwork", "RedunEdge", "gene_network_from_theta", "hill_activation",
    "hill_repression", "is_bursty", "n_species", "precompute_hill_matrices",
    "compare_moments", "convergence_rate", "ks_critical_value", "ks_statistic",
]
