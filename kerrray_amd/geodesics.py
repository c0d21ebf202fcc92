"""Integrator options, state layout and the Runge-Kutta tableaus (host side).

The option dataclasses carry the names, defaults and validation of the
reference (src/kerrray/geodesics/integrators.py:67-162); the tableaus are
the Dormand-Prince 5(4) and classical RK4 coefficients of
src/kerrray/geodesics/tableaus.py:59-91 (Dormand & Prince 1980; Hairer,
Norsett & Wanner 1993, Tables 1.2 and 5.2), evaluated in float64 exactly as
the reference evaluates them, because the device kernels are checked
against these values bit for bit before every launch.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

__all__ = [
    "DP_A", "DP_B", "DP_C", "DP_E", "ERROR_EXPONENT", "EventOptions", "FACTOR_MAX", "FACTOR_MIN",
    "IDX_PH", "IDX_PPH", "IDX_PR", "IDX_PT", "IDX_PTH", "IDX_R", "IDX_T", "IDX_TH", "IntegratorOptions",
    "RADICAND_TOLERANCE", "RK4_B", "RK4_C", "SAFETY", "STATE_SIZE", "SUPPORTED_METHODS", "TerminationOptions",
    "carter_potentials", "max_attempts", "photon_from_constants",
]

IDX_T, IDX_R, IDX_TH, IDX_PH, IDX_PT, IDX_PR, IDX_PTH, IDX_PPH = range(8)
STATE_SIZE = 8

RK4_C = (0.0, 0.5, 0.5, 1.0)
RK4_B = (1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0)
DP_C = (0.0, 1.0 / 5.0, 3.0 / 10.0, 4.0 / 5.0, 8.0 / 9.0, 1.0, 1.0)
DP_A = (
    (),
    (1.0 / 5.0,),
    (3.0 / 40.0, 9.0 / 40.0),
    (44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0),
    (19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0),
    (9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0),
    (35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0),
)
DP_B = (35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0, 0.0)
DP_B_HAT = (5179.0 / 57600.0, 0.0, 7571.0 / 16695.0, 393.0 / 640.0, -92097.0 / 339200.0, 187.0 / 2100.0, 1.0 / 40.0)
DP_E = tuple(bh - b for bh, b in zip(DP_B_HAT, DP_B))
SAFETY = 0.9
FACTOR_MIN = 0.2
FACTOR_MAX = 5.0
ERROR_EXPONENT = -1.0 / 5.0

SUPPORTED_METHODS = frozenset({"rk4", "rk45", "dop853"})


@dataclass(frozen=True)
class IntegratorOptions:
    """Scheme, tolerances and budgets (integrators.py:70-115)."""

    method: str = "rk45"
    rtol: float = 1e-9
    atol: float = 1e-11
    step_size: float = 0.01
    max_steps: int = 100_000
    lambda_max: float = 1.0e4
    h_min: float = 1e-12
    dtype: str = "float64"
    h_max: float | None = None

    def __post_init__(self) -> None:
        if self.method not in SUPPORTED_METHODS:
            raise ValueError(f"method must be one of {sorted(SUPPORTED_METHODS)}, got {self.method!r}")
        for name in ("rtol", "atol", "step_size", "lambda_max", "h_min"):
            value = getattr(self, name)
            if not (math.isfinite(value) and value > 0.0):
                raise ValueError(f"{name} must be a finite number > 0, got {value!r}")
        if self.h_max is not None and not (math.isfinite(self.h_max) and self.h_max > 0.0):
            raise ValueError(f"h_max must be None or a finite number > 0, got {self.h_max!r}")
        if self.max_steps < 1:
            raise ValueError(f"max_steps must be >= 1, got {self.max_steps!r}")
        if self.dtype not in ("float32", "float64"):
            raise ValueError(f"dtype must be 'float32' or 'float64', got {self.dtype!r}")

    @property
    def np_dtype(self) -> np.dtype:
        return np.dtype(self.dtype)


@dataclass(frozen=True)
class TerminationOptions:
    """Capture margin and escape radius (integrators.py:118-129)."""

    horizon_epsilon: float = 1e-6
    escape_radius: float = 1000.0

    def __post_init__(self) -> None:
        if not (math.isfinite(self.horizon_epsilon) and self.horizon_epsilon >= 0.0):
            raise ValueError(f"horizon_epsilon must be >= 0, got {self.horizon_epsilon!r}")
        if not (math.isfinite(self.escape_radius) and self.escape_radius > 0.0):
            raise ValueError(f"escape_radius must be > 0, got {self.escape_radius!r}")


@dataclass(frozen=True)
class EventOptions:
    """Optional equatorial disk-plane event (integrators.py:132-148)."""

    disk_plane: bool = False
    r_in: float = 6.0
    r_out: float = 20.0

    def __post_init__(self) -> None:
        if not (0.0 < self.r_in <= self.r_out):
            raise ValueError(f"need 0 < r_in <= r_out, got r_in={self.r_in!r}, r_out={self.r_out!r}")


def max_attempts(integ: IntegratorOptions) -> int:
    """Defensive bound on step attempts before ``h < h_min`` (integrators.py:151-162)."""
    h_max = max(integ.lambda_max, integ.step_size)
    rej = (math.log(h_max / integ.h_min) + integ.max_steps * math.log(FACTOR_MAX)) / math.log(1.0 / SAFETY)
    return integ.max_steps + int(math.ceil(rej)) + 1


RADICAND_TOLERANCE = 1e-12
"""Radicands ``R`` or ``Theta`` more negative than this times their scale raise (initial_conditions.py:46)."""


def carter_potentials(st, r: float, theta: float, E: float, Lz: float, Q: float) -> tuple[float, float, float, float]:
    """``(R(r), Theta(theta), scale_R, scale_Theta)`` for the constants ``(E, L_z, Q)`` (initial_conditions.py:50-68)."""
    a = st.a
    dlt = r * r - 2.0 * st.mass * r + a * a
    term1 = E * (r * r + a * a) - a * Lz
    term2 = dlt * (Q + (Lz - a * E) ** 2)
    big_r = term1 * term1 - term2
    scale_r = term1 * term1 + abs(term2)
    c2, s2 = math.cos(theta) ** 2, math.sin(theta) ** 2
    lz_term = 0.0 if Lz == 0.0 else Lz * Lz / s2
    big_theta = Q - c2 * (lz_term - a * a * E * E)
    scale_theta = abs(Q) + c2 * (lz_term + a * a * E * E)
    return big_r, big_theta, scale_r, scale_theta


def photon_from_constants(st, r: float, theta: float, E: float, Lz: float, Q: float, *, sign_r: int,
                          sign_theta: int, t: float = 0.0, phi: float = 0.0) -> np.ndarray:
    """Photon state at ``(t, r, theta, phi)`` with constants ``(E, L_z, Q)`` (initial_conditions.py:83-112).

    ``p_r = sign_r sqrt(R) / Delta``, ``p_theta = sign_theta sqrt(Theta)``;
    ``ValueError`` for ``E <= 0``, a point not outside the horizon, a sign
    other than +-1, or a radicand negative beyond round-off.
    """
    from kerrray_amd.geometry import outer_horizon

    if not E > 0.0:
        raise ValueError(f"E must be > 0 for a future-directed photon, got {E!r}")
    if not r > outer_horizon(st):
        raise ValueError(f"r = {r!r} must lie outside the outer horizon r_plus = {outer_horizon(st)!r}")
    for sign, name in ((sign_r, "sign_r"), (sign_theta, "sign_theta")):
        if sign not in (-1, 1):
            raise ValueError(f"{name} must be +1 or -1, got {sign!r}")
    big_r, big_theta, scale_r, scale_theta = carter_potentials(st, r, theta, E, Lz, Q)
    for value, scale, name in ((big_r, scale_r, "R(r)"), (big_theta, scale_theta, "Theta(theta)")):
        if value < -RADICAND_TOLERANCE * max(scale, 1.0):
            raise ValueError(f"{name} = {value!r} < 0: the requested constants of motion are inconsistent here")
    dlt = r * r - 2.0 * st.mass * r + st.a**2
    p_r = float(sign_r) * math.sqrt(max(big_r, 0.0)) / dlt
    p_theta = float(sign_theta) * math.sqrt(max(big_theta, 0.0))
    return np.array([t, r, theta, phi, -E, p_r, p_theta, Lz], dtype=np.float64)
