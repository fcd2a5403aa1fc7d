"""Kinetic, population and capture parameter types.

KineticParams and steady_state follow src/kinetics.jl:25-70;
PopulationConfig and CaptureModel follow src/population.jl:27-74.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib


@dataclass(frozen=True)
class KineticParams:
    """k_t, K_d, n, mu_m, mu_p, dilution with the reference defaults (kinetics.jl:41-50)."""
    k_t: float = 4.3376521035785345
    K_d: float = 61.47011276535837
    n: float = 10.0
    mu_m: float = 0.04987442692380366
    mu_p: float = 0.36777724535966944
    dilution: float = 0.0

    def to_c(self) -> _lib.ScrnaKinetics:
        return _lib.ScrnaKinetics(self.k_t, self.K_d, self.n, self.mu_m, self.mu_p, self.dilution)


def steady_state(net, kin: KineticParams):
    """m_ss = beta/(mu_m + mu), p_ss = k_t m_ss/(mu_p + mu) (kinetics.jl:65-70)."""
    mu = kin.dilution
    m_ss = np.asarray(net.basals, dtype=np.float64) / (kin.mu_m + mu)
    p_ss = kin.k_t * m_ss / (kin.mu_p + mu)
    return m_ss, p_ss


@dataclass(frozen=True)
class PopulationConfig:
    """Moran fixed-N population (population.jl:27-45).

    ``partition`` and ``replace`` choose between the reference's two division
    rules: "binomial" partitioning with replacement over all N cells is
    population.jl:157-205; "halve" with "exclude_mother" is the CPU CLE
    (cle.jl:249-272); "halve" with "uniform" is the CUDA extension
    (ext/SyntheticscRNAseqCUDAExt.jl:494-513).  ``growth_noise`` (population.jl:128-139)
    randomises each cell's volume growth in the binomial tau-leap; the CLE grows
    volumes deterministically and ignores it, as cle.jl:169 does.
    """
    cell_num: int = 1000
    growth_rate: float = 0.03
    V_div: float = 2.0
    V_init: tuple = (0.8, 1.2)
    div_check_interval: int = 10
    growth_noise: float = 0.0
    partition: str = "binomial"
    replace: str = "uniform"

    def to_c(self) -> _lib.ScrnaPopulation:
        part = {"halve": 0, "binomial": 1}[self.partition]
        rep = {"uniform": 0, "exclude_mother": 1}[self.replace]
        return _lib.ScrnaPopulation(1, self.growth_rate, self.V_div, float(self.V_init[0]),
                                    float(self.V_init[1]), int(self.div_check_interval), part, rep,
                                    float(self.growth_noise))


@dataclass(frozen=True)
class CaptureModel:
    """LogNormal-Binomial capture (population.jl:57-74); readout is :mrna or :protein."""
    efficiency: float = 0.1
    efficiency_std: float = 0.2
    readout: str = "mrna"

    def __post_init__(self):
        if self.readout not in ("mrna", "protein"):            # population.jl:63
            raise ValueError("readout must be 'mrna' or 'protein'")

    def to_c(self) -> _lib.ScrnaCapture:
        return _lib.ScrnaCapture(1, self.efficiency, self.efficiency_std, _lib.READOUT[self.readout])
