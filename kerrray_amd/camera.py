"""Image-plane camera with Bardeen initial conditions.

:class:`BardeenCamera` has the fields and pixel grid of the reference's
``Camera`` (src/kerrray/raytracing/camera.py:100-222): ``resolution`` x
``resolution`` pixel centres ``alpha_j = fov (2 j + 1 - n) / n``,
``beta_i = -alpha_i``, row 0 on top. Each pixel's photon is built from the
celestial constants of its image point, as the reference's
``_off_equatorial_state`` does (src/kerrray/benchmarks/rayset.py:158-165,
src/kerrray/geodesics/initial_conditions.py:50-112)::

    xi  = -alpha sin(theta_o)                       (= L_z / E)
    eta = beta^2 - a^2 cos^2(theta_o) + xi^2 cot^2(theta_o)   (= Q / E^2)
    y0  = [0, r_o, theta_o, 0, -E, -sqrt(R(r_o))/Delta, s_theta sqrt(Theta(theta_o)), xi],  E = 1

(forward in lambda, moving inward). ``s_theta = -1`` for ``beta >= 0`` and
``+1`` otherwise, the direction in which the backward ray of a pixel above
the image centre leaves the observer. Capture depends only on
``(xi, eta)``, so the mask does not depend on ``s_theta``.

The device kernels generate these states themselves (one pixel per lane);
:func:`bardeen_initial_states` builds the same states on the host, e.g. to
feed :meth:`Backend.integrate_batch`.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
from numpy.typing import ArrayLike, NDArray

from kerrray_amd.geometry import Spacetime, outer_horizon

__all__ = ["INCLINATION_EPSILON_DEG", "RADICAND_TOLERANCE", "BardeenCamera", "bardeen_initial_states", "bardeen_points"]

INCLINATION_EPSILON_DEG = 1e-3
RADICAND_TOLERANCE = 1e-12


@dataclass(frozen=True)
class BardeenCamera:
    """Observer at ``(radius, inclination_deg)`` with a square image plane of half-width ``fov``."""

    radius: float
    inclination_deg: float
    phi_deg: float = 0.0
    fov: float = 12.0
    resolution: int = 64

    def __post_init__(self) -> None:
        radius = float(self.radius)
        inclination = float(self.inclination_deg)
        fov = float(self.fov)
        if not (math.isfinite(radius) and radius > 0.0):
            raise ValueError(f"radius must be a finite number > 0, got {self.radius!r}")
        if not (math.isfinite(inclination) and 0.0 <= inclination <= 180.0):
            raise ValueError(f"inclination_deg must lie in [0, 180], got {self.inclination_deg!r}")
        if float(self.phi_deg) != 0.0:
            raise ValueError("the Bardeen camera sits at phi_o = 0 (rayset.py:158-165); got phi_deg != 0")
        if not (math.isfinite(fov) and fov > 0.0):
            raise ValueError(f"fov must be a finite number > 0, got {self.fov!r}")
        res = self.resolution
        if isinstance(res, bool) or not isinstance(res, (int, np.integer)):
            raise ValueError(f"resolution must be an integer, got {res!r}")
        res = int(res)
        if res < 2 or res % 2 != 0:
            raise ValueError(f"resolution must be an even integer >= 2, got {res!r}")
        low, high = INCLINATION_EPSILON_DEG, 180.0 - INCLINATION_EPSILON_DEG
        object.__setattr__(self, "radius", radius)
        object.__setattr__(self, "inclination_deg", min(max(inclination, low), high))
        object.__setattr__(self, "phi_deg", 0.0)
        object.__setattr__(self, "fov", fov)
        object.__setattr__(self, "resolution", res)

    @property
    def n_rays(self) -> int:
        return self.resolution * self.resolution

    @property
    def shape(self) -> tuple[int, int]:
        return (self.resolution, self.resolution)

    @property
    def pixel_size(self) -> float:
        return 2.0 * self.fov / self.resolution

    @property
    def inclination_rad(self) -> float:
        return math.radians(self.inclination_deg)

    @property
    def phi_rad(self) -> float:
        return 0.0

    def pixel_coordinates(self) -> tuple[NDArray[np.float64], NDArray[np.float64]]:
        """``(alpha, beta)`` of the pixel centres, shape ``(H, W)`` (camera.py:206-222)."""
        n = self.resolution
        centres = self.fov * (2.0 * np.arange(n, dtype=np.float64) + 1.0 - n) / n
        alpha, beta = np.meshgrid(centres, -centres)
        return alpha, beta


def bardeen_points(
    st: Spacetime, r_o: float, inclination_deg: float, alpha: ArrayLike, beta: ArrayLike
) -> NDArray[np.float64]:
    """Bardeen photons (N, 8) for image points ``(alpha, beta)`` (module docstring).

    Vectorised over the points with the reference's scalar expressions;
    NumPy's elementwise ``+ - * / sqrt`` round like Python floats, and the
    per-image factors use Python's own ``math`` and ``**``. Raises
    ``ValueError`` where ``photon_from_constants`` raises.
    """
    if not r_o > outer_horizon(st):
        raise ValueError(f"r = {r_o!r} must lie outside the outer horizon r_plus = {outer_horizon(st)!r}")
    th = math.radians(inclination_deg)
    sin_o, cos_o = math.sin(th), math.cos(th)
    a, m = st.a, st.mass
    alpha = np.asarray(alpha, dtype=np.float64).ravel()
    beta = np.asarray(beta, dtype=np.float64).ravel()
    a2c2 = a**2 * cos_o**2
    cot2 = (cos_o / sin_o) ** 2
    xi = -alpha * sin_o
    eta = beta * beta - a2c2 + xi * xi * cot2
    dlt = r_o * r_o - 2.0 * m * r_o + a**2
    E = 1.0
    term1 = E * (r_o * r_o + a * a) - a * xi
    d = xi - a * E
    term2 = dlt * (eta + d * d)
    big_r = term1 * term1 - term2
    scale_r = term1 * term1 + np.abs(term2)
    c2, s2 = math.cos(th) ** 2, math.sin(th) ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        lz_term = np.where(xi == 0.0, 0.0, xi * xi / s2)
    aaee = a * a * E * E
    big_theta = eta - c2 * (lz_term - aaee)
    scale_theta = np.abs(eta) + c2 * (lz_term + aaee)
    if np.any(big_r < -RADICAND_TOLERANCE * np.maximum(scale_r, 1.0)):
        raise ValueError("R(r) < 0: the requested constants of motion are inconsistent here")
    if np.any(big_theta < -RADICAND_TOLERANCE * np.maximum(scale_theta, 1.0)):
        raise ValueError("Theta(theta) < 0: the requested constants of motion are inconsistent here")
    y = np.empty((alpha.size, 8))
    y[:, 0] = 0.0
    y[:, 1] = r_o
    y[:, 2] = th
    y[:, 3] = 0.0
    y[:, 4] = -E
    y[:, 5] = -1.0 * np.sqrt(np.maximum(big_r, 0.0)) / dlt
    y[:, 6] = np.where(beta >= 0.0, -1.0, 1.0) * np.sqrt(np.maximum(big_theta, 0.0))
    y[:, 7] = xi
    return y


def bardeen_initial_states(cam: BardeenCamera, st: Spacetime) -> NDArray[np.float64]:
    """Row-major (n*n, 8) initial states of every pixel of ``cam``."""
    alpha, beta = cam.pixel_coordinates()
    return bardeen_points(st, cam.radius, cam.inclination_deg, alpha, beta)
