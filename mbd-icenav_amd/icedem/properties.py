"""Ice and ship properties (``medea_ice_ship/properties.py:6-54``)."""
from __future__ import annotations

import numpy as np


class IceProperties:
    ICE_DENSITY = 920
    WATER_DENSITY = 1025
    FLEXURAL_STRENGTH_BASE = 0.3e6
    YOUNGS_MODULUS = 5e9

    @staticmethod
    def flexural_strength(thickness, base=FLEXURAL_STRENGTH_BASE):
        return base * np.sqrt(thickness / 1.0)

    @staticmethod
    def breaking_force(thickness, width, base=FLEXURAL_STRENGTH_BASE):
        sigma_f = IceProperties.flexural_strength(thickness, base)
        return sigma_f * width * thickness ** 2


class ShipProperties:
    """Mass = displacement = L B T 0.6 rho_w unless given (properties.py:25-38)."""

    def __init__(self, length, beam, draft, displacement=None):
        self.length = length
        self.beam = beam
        self.draft = draft
        if displacement is None:
            self.displacement = length * beam * draft * 0.6 * IceProperties.WATER_DENSITY
        else:
            self.displacement = displacement
        self.mass = self.displacement


def ssa(angle):
    """Smallest signed angle in [-pi, pi) (properties.py:51-54)."""
    return (angle + np.pi) % (2 * np.pi) - np.pi
