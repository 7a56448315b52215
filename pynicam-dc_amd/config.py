"""Physical constants and run parameters of the dynamical-core step.

Defaults follow the reference constant table (share/mod_const.py:9-205) with
the overrides the reference test case applies in its ``[cnstparam]`` section
(pynicamdc/test/case2/config/nhm_driver.toml: gravity 9.80616, R 287.0,
Cp 1004.5), the sphere radius of the benchmark (6.371e6 m) and the numerical
filter / time-integration settings of the benchmark family
(RK3, 6 small steps per large step, DIRECT hyperdiffusion and 3-D divergence
damping of order 2).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Const:
    RADIUS: float = 6.371e6
    OHM: float = 7.292e-5
    GRAV: float = 9.80616
    Rdry: float = 287.0
    CPdry: float = 1004.5
    Rvap: float = 461.46
    CPvap: float = 1845.60
    PRE00: float = 1.0e5
    UNDEF: float = -9.9999e30
    ALPHA: float = 1.0                 # non-hydrostatic switch (mod_runconf.py:13)
    # tracer layout of a dry run: one water tracer (qv), mod_runconf.py:191-227
    NQW_STR: int = 0
    NQW_END: int = 0
    I_QV: int = 0
    NTRACER: int = 1
    # numerical filters (mod_numfilter.py:51-217)
    GAMMA_H: float = 1.0e16            # DIRECT hyperdiffusion coefficient [m^4/s]
    LAP_ORDER_HDIFF: int = 2
    HDIFF_FACT_RHO: float = 1.0e-2
    ALPHA_D: float = 1.0e16            # DIRECT 3-D divergence damping coefficient
    LAP_ORDER_DIVDAMP: int = 2
    # latent heat of vaporisation for the energy budget (mod_const.py:53-81, 183-192)
    LHV0: float = 2.5008e6
    CL: float = 4218.0
    TEM00: float = 273.15
    THERMODYN_TYPE: str = "SIMPLE"
    # tracer transport (mod_runconf.py:19, TRC_ADV_TYPE = MIURA2004)
    THUBURN_LIM: bool = True
    # time integration (mod_time.py:83-102)
    DTL: float = 1200.0
    SSTEP_MAX: int = 6
    CVW: np.ndarray = field(default=None)

    def __post_init__(self):
        if self.CVW is None:
            self.CVW = np.array([self.CVvap] + [0.0] * (self.NTRACER - 1))

    @property
    def CVdry(self) -> float:
        return self.CPdry - self.Rdry

    @property
    def CVvap(self) -> float:
        return self.CPvap - self.Rvap

    @property
    def LHV(self) -> float:
        """LHV0 for the SIMPLE energy type, LHV00 = LHV0 - (CPvap - CL) TEM00 for EXACT."""
        if self.THERMODYN_TYPE == "EXACT":
            return self.LHV0 - (self.CPvap - self.CL) * self.TEM00
        if self.THERMODYN_TYPE in ("SIMPLE", "SIMPLE2"):
            return self.LHV0
        raise ValueError(f"unknown thermodynamics type {self.THERMODYN_TYPE!r}")

    @property
    def DTS(self) -> float:
        return self.DTL / float(self.SSTEP_MAX)


def benchmark_const(glevel: int, dtl: float) -> Const:
    """Constants for the benchmark family at one resolution.

    The hyperdiffusion coefficient is 1e16 m^4/s at glevel 5 and is divided by
    8 per finer level (factor 2^3), the resolution scaling the reference uses
    in its sweep table (tools/sweep/README.md:39-52).
    """
    gamma = 1.0e16 / (8.0 ** max(glevel - 5, 0))
    return Const(GAMMA_H=gamma, ALPHA_D=gamma, DTL=dtl)
