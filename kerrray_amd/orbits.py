"""Spherical photon orbits, critical impact parameters and Bardeen's shadow curve.

Host-side analytic references of the reference (src/kerrray/photons/orbits.py:103-273),
used to bracket the edge search and to score it. Formulas:

* spherical photon orbit constants (R = 0 = dR/dr; Bardeen 1973; Teo 2003)::

      xi(r)  = -[r^2 (r - 3M) + a^2 (r + M)] / [a (r - M)]
      eta(r) = r^3 [4 M a^2 - r (r - 3M)^2] / [a^2 (r - M)^2]

* equatorial circular photon orbits: r - 3M = -+ 2|a| sqrt(M / r) (upper sign
  prograde), closed form BPT 1972 eq. 2.18 r = 2M {1 + cos[(2/3) arccos(-+|a|/M)]};
* Bardeen's curve for an observer at inclination i:
  alpha = -xi / sin i, beta = +- sqrt(eta + a^2 cos^2 i - xi^2 cot^2 i),
  r between the roots of the radicand on either side of the polar orbit
  (xi = 0); for a = 0 the circle of radius 3 sqrt(3) M.

For initial conditions built from (xi, eta) (Bardeen image-plane
coordinates) the captured region of an image is exactly the inside of this
curve, at any observer radius outside the photon region.
"""

from __future__ import annotations

import math

import numpy as np
from scipy.optimize import brentq

from kerrray_amd.geometry import Spacetime

__all__ = [
    "critical_impact_parameters",
    "equatorial_photon_orbit_radius",
    "equatorial_photon_orbit_radius_closed_form",
    "shadow_curve",
    "spherical_orbit_constants",
]

_XTOL = 1e-14
_RTOL = 4 * np.finfo(float).eps


def spherical_orbit_constants(st: Spacetime, r) -> tuple[np.ndarray, np.ndarray]:
    """``(xi, eta)`` of the spherical photon orbit at radius ``r`` (vectorised; a != 0)."""
    m, a = st.mass, st.a
    if a == 0.0:
        raise ValueError("spherical_orbit_constants is undefined for a = 0 (all orbits at r = 3M)")
    r = np.asarray(r, dtype=np.float64)
    r_m3 = r - 3.0 * m
    xi = -(r * r * r_m3 + a * a * (r + m)) / (a * (r - m))
    eta = r**3 * (4.0 * m * a * a - r * r_m3 * r_m3) / (a * a * (r - m) ** 2)
    return xi, eta


def equatorial_photon_orbit_radius(st: Spacetime, prograde: bool = True) -> float:
    """Equatorial circular photon orbit radius by bracketed root finding (regular form)."""
    m, abs_a = st.mass, abs(st.a)
    s = 1.0 if prograde else -1.0

    def f(r: float) -> float:
        return r - 3.0 * m + s * 2.0 * abs_a * math.sqrt(m / r)

    lo, hi = (m, 3.5 * m) if prograde else (2.5 * m, 4.0 * m)
    return float(brentq(f, lo, hi, xtol=_XTOL * m, rtol=_RTOL))


def equatorial_photon_orbit_radius_closed_form(st: Spacetime, prograde: bool = True) -> float:
    """BPT 1972 eq. 2.18."""
    s = 1.0 if prograde else -1.0
    return 2.0 * st.mass * (1.0 + math.cos((2.0 / 3.0) * math.acos(-s * abs(st.a) / st.mass)))


def critical_impact_parameters(st: Spacetime) -> tuple[float, float]:
    """``(b_prograde, b_retrograde) = |xi|`` on the equatorial photon orbits (3 sqrt 3 M at a = 0)."""
    m, a = st.mass, st.a
    out = []
    for prograde in (True, False):
        r = equatorial_photon_orbit_radius(st, prograde)
        s = 1.0 if prograde else -1.0
        delta = r * r - 2.0 * m * r + a * a
        sign_a = 1.0 if a >= 0.0 else -1.0
        # eta + (xi - a)^2 = 4 r^2 Delta / (r - M)^2 with eta = 0 on the equator.
        out.append(abs(sign_a * (abs(a) + s * 2.0 * r * math.sqrt(delta) / (r - m))))
    return out[0], out[1]


def shadow_curve(st: Spacetime, inclination_deg: float, n: int = 720) -> tuple[np.ndarray, np.ndarray]:
    """Closed curve ``(alpha, beta)`` of ``2 (n // 2)`` points ordered by polar angle about the origin."""
    m, a = st.mass, st.a
    if not 0.0 < inclination_deg < 180.0:
        raise ValueError("inclination_deg must lie strictly between 0 and 180")
    n_points = 2 * (int(n) // 2)
    if n_points < 4:
        raise ValueError("n must be at least 4")
    inc = math.radians(inclination_deg)
    sin_i = math.sin(inc)
    cos_i = 0.0 if inclination_deg == 90.0 else math.cos(inc)
    if a == 0.0:
        b_c = critical_impact_parameters(st)[0]
        ang = np.linspace(-math.pi, math.pi, n_points, endpoint=False)
        return b_c * np.cos(ang), b_c * np.sin(ang)
    cot2 = (cos_i / sin_i) ** 2
    r_lo = equatorial_photon_orbit_radius(st, True)
    r_hi = equatorial_photon_orbit_radius(st, False)

    def xi_of(r: float) -> float:
        return float(spherical_orbit_constants(st, r)[0])

    def radicand(r: float) -> float:
        xi, eta = spherical_orbit_constants(st, r)
        return float(eta + a * a * cos_i * cos_i - xi * xi * cot2)

    r0 = brentq(xi_of, r_lo, r_hi, xtol=_XTOL * m, rtol=_RTOL)
    if radicand(r0) <= 0.0:
        raise RuntimeError("radicand is not positive at the polar orbit")
    r_a = r_lo if radicand(r_lo) >= 0.0 else brentq(radicand, r_lo, r0, xtol=_XTOL * m, rtol=_RTOL)
    r_b = r_hi if radicand(r_hi) >= 0.0 else brentq(radicand, r0, r_hi, xtol=_XTOL * m, rtol=_RTOL)
    k = n_points // 2 + 1
    u = np.linspace(0.0, 1.0, k)
    r = r_a + (r_b - r_a) * 0.5 * (1.0 - np.cos(math.pi * u))  # cluster at the turning points
    xi, eta = spherical_orbit_constants(st, r)
    beta_up = np.sqrt(np.clip(eta + a * a * cos_i * cos_i - xi * xi * cot2, 0.0, None))
    beta_up[0] = 0.0
    beta_up[-1] = 0.0
    alpha_up = -xi / sin_i
    alpha = np.concatenate([alpha_up, alpha_up[-2:0:-1]])
    beta = np.concatenate([beta_up, -beta_up[-2:0:-1]])
    order = np.argsort(np.arctan2(beta, alpha), kind="stable")
    return alpha[order], beta[order]


def isco_radius(st: Spacetime, prograde: bool = True) -> float:
    """Marginally stable circular orbit radius, BPT 1972 eq. 2.21 (orbits.py:336-350).

    Z1 = 1 + (1 - chi^2)^(1/3) [(1 + chi)^(1/3) + (1 - chi)^(1/3)], Z2 = (3 chi^2 + Z1^2)^(1/2),
    r_ms = M {3 + Z2 -+ [(3 - Z1)(3 + Z1 + 2 Z2)]^(1/2)}, chi = |a| / M, upper sign prograde.
    """
    m = st.mass
    s = 1.0 if prograde else -1.0
    chi = abs(st.a) / m
    z1 = 1.0 + (1.0 - chi * chi) ** (1.0 / 3.0) * ((1.0 + chi) ** (1.0 / 3.0) + (1.0 - chi) ** (1.0 / 3.0))
    z2 = math.sqrt(3.0 * chi * chi + z1 * z1)
    return m * (3.0 + z2 - s * math.sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2)))
