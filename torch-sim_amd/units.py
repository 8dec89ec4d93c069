"""Metal unit system (Angstrom, eV, amu, e).

Same constants and derived units as torch_sim/units.py:38-114 (CODATA 2014
values; the conversion factors must be identical for drop-in unit handling
of ``integrate``'s ``timestep`` in ps and ``temperature`` in K).
"""

from __future__ import annotations

import math
from enum import Enum


class BaseConstant(float, Enum):
    """CODATA 2014 constants (torch_sim/units.py:38-62)."""

    e = 1.6021766208e-19  # elementary charge, C
    amu = 1.660539040e-27  # atomic mass unit, kg
    k_B = 1.38064852e-23  # Boltzmann constant, J/K


class MetalUnits(float, Enum):
    """Derived metal units (torch_sim/units.py:94-114)."""

    mass = 1.0
    distance = 1.0
    energy = 1.0
    charge = 1.0
    # internal time units per ps (~98.2269): dt_internal = dt_ps * time
    time = math.sqrt(1.0 * BaseConstant.e / (BaseConstant.amu * 1e-20)) * 1e-12
    velocity = 1.0 / time
    force = 1.0
    temperature = BaseConstant.k_B / BaseConstant.e  # eV / K
    pressure = 1e5 * (1.0 * 1e-30 / BaseConstant.e)  # bar in eV / A^3
