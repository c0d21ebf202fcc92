"""Kerr spacetime parameters and horizons (host side).

Mirrors ``Spacetime`` and the horizon helpers of the reference
(src/kerrray/geometry/metric.py:66-98, src/kerrray/geometry/horizons.py:33-42):
same fields, same validation, same floating-point expressions.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

__all__ = ["Spacetime", "horizon_radii", "kerr", "outer_horizon", "schwarzschild"]


@dataclass(frozen=True)
class Spacetime:
    """A Kerr black hole in geometric units: ``mass > 0`` and dimensionless ``|spin| < 1``."""

    mass: float = 1.0
    spin: float = 0.0

    def __post_init__(self) -> None:
        mass = float(self.mass)
        spin = float(self.spin)
        if not (math.isfinite(mass) and mass > 0.0):
            raise ValueError(f"mass must be a finite number > 0, got {self.mass!r}")
        if not (math.isfinite(spin) and abs(spin) < 1.0):
            raise ValueError(f"spin must satisfy abs(spin) < 1, got {self.spin!r}")
        object.__setattr__(self, "mass", mass)
        object.__setattr__(self, "spin", spin)

    @property
    def a(self) -> float:
        """Kerr parameter ``a = spin * mass``."""
        return self.spin * self.mass

    @property
    def is_schwarzschild(self) -> bool:
        return self.spin == 0.0


def schwarzschild(mass: float = 1.0) -> Spacetime:
    return Spacetime(mass=mass, spin=0.0)


def kerr(mass: float, spin: float) -> Spacetime:
    return Spacetime(mass=mass, spin=spin)


def horizon_radii(st: Spacetime) -> tuple[float, float]:
    """``(r_plus, r_minus) = M +- sqrt(M^2 - a^2)``."""
    m = st.mass
    root = math.sqrt(m * m - st.a * st.a)
    return m + root, m - root


def outer_horizon(st: Spacetime) -> float:
    """The outer horizon ``r_plus``, the capture radius."""
    return horizon_radii(st)[0]
