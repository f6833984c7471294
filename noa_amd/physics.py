"""PMS scalar types and constants (src/noa/pms/physics.hh:29-80)."""
from __future__ import annotations

from dataclasses import dataclass

AVOGADRO_NUMBER = 6.02214076E+23      # physics.hh:54
ATOMIC_MASS_ENERGY = 0.931494         # physics.hh:55
ELECTRON_MASS = 0.510998910E-03       # GeV/c^2, physics.hh:57
MUON_MASS = 0.10565839                # GeV/c^2, physics.hh:58
TAU_MASS = 1.77682                    # GeV/c^2, physics.hh:59
MUON_CTAU = 658.654                   # m, physics.hh:61
TAU_CTAU = 87.03E-06                  # m, physics.hh:62
LARMOR_FACTOR = 0.299792458           # physics.hh:64
X_FRACTION = 5E-02                    # physics.hh:76
KIN_CUTOFF = 1E-9                     # physics.hh:78


@dataclass(frozen=True)
class AtomicElement:
    """physics.hh:39-43: atomic mass A (g/mol), mean excitation I (GeV), Z."""
    A: float
    I: float
    Z: int


STANDARD_ROCK = AtomicElement(A=22.0, I=0.1364E-6, Z=11)   # physics.hh:67-71
