"""Analytic references used to bracket and score the edge: closed forms and invariants."""

from __future__ import annotations

import math

import numpy as np
import pytest

from kerrray_amd.bisection import rounds_for_width
from kerrray_amd.boundary import polygon_centroid, polygon_radii
from kerrray_amd.camera import BardeenCamera
from kerrray_amd.geometry import Spacetime
from kerrray_amd.orbits import (
    critical_impact_parameters, equatorial_photon_orbit_radius, equatorial_photon_orbit_radius_closed_form,
    shadow_curve, spherical_orbit_constants,
)


def test_schwarzschild_critical_impact_parameter_is_3_sqrt3() -> None:
    b_pro, b_ret = critical_impact_parameters(Spacetime(1.0, 0.0))
    assert b_pro == pytest.approx(3.0 * math.sqrt(3.0), rel=1e-12) and b_ret == pytest.approx(b_pro, rel=1e-12)


@pytest.mark.parametrize("spin", [0.3, 0.9, 0.99])
def test_equatorial_orbits_match_bpt_closed_form(spin: float) -> None:
    st = Spacetime(1.0, spin)
    for prograde in (True, False):
        assert equatorial_photon_orbit_radius(st, prograde) == pytest.approx(
            equatorial_photon_orbit_radius_closed_form(st, prograde), rel=1e-12)


@pytest.mark.parametrize("spin", [0.5, 0.9])
def test_spherical_orbits_are_double_roots_of_R(spin: float) -> None:
    """R(r) = (r^2 + a^2 - a xi)^2 - Delta (eta + (xi - a)^2) and dR/dr vanish on the orbit."""
    st = Spacetime(1.0, spin)
    a = st.a
    for r in np.linspace(equatorial_photon_orbit_radius(st, True), equatorial_photon_orbit_radius(st, False), 7):
        xi, eta = (float(v) for v in spherical_orbit_constants(st, r))

        def R(x):
            return (x * x + a * a - a * xi) ** 2 - (x * x - 2 * x + a * a) * (eta + (xi - a) ** 2)

        h = 1e-5
        assert abs(R(r)) < 1e-8 * r**4
        assert abs((R(r + h) - R(r - h)) / (2 * h)) < 1e-6 * r**3


def test_shadow_curve_properties() -> None:
    st = Spacetime(1.0, 0.9)
    a, b = shadow_curve(st, 60.0, 720)
    assert a.size == 720 and np.all(np.diff(np.arctan2(b, a)) >= 0)
    # beta -> -beta symmetric and displaced towards alpha > 0 for a > 0.
    assert np.allclose(np.sort(b), np.sort(-b), atol=1e-12)
    assert polygon_centroid(a, b)[0] > 0.0
    a0, b0 = shadow_curve(Spacetime(1.0, 0.0), 60.0, 360)
    assert np.allclose(np.hypot(a0, b0), 3.0 * math.sqrt(3.0), rtol=1e-12)


def test_polygon_radii_of_a_circle() -> None:
    ang = np.linspace(0, 2 * math.pi, 4000, endpoint=False)
    x, y = 2.0 + 3.0 * np.cos(ang), -1.0 + 3.0 * np.sin(ang)
    r = polygon_radii(x, y, (2.0, -1.0), np.array([0.1, 1.0, 2.0, 4.0]))
    assert np.allclose(r, 3.0, rtol=1e-6)
    cx, cy = polygon_centroid(x, y)
    assert abs(cx - 2.0) < 1e-12 and abs(cy + 1.0) < 1e-12


def _brute_radii(x, y, centre, angles):
    """Smallest valid ray parameter over every segment: the definition polygon_radii evaluates."""
    qx, qy = x - centre[0], y - centre[1]
    ex, ey = np.roll(qx, -1) - qx, np.roll(qy, -1) - qy
    dx, dy = np.cos(angles)[:, None], np.sin(angles)[:, None]
    det = dx * (-ey[None, :]) - dy * (-ex[None, :])
    with np.errstate(divide="ignore", invalid="ignore"):
        t = (qx[None, :] * (-ey[None, :]) - qy[None, :] * (-ex[None, :])) / det
        s = (dx * qy[None, :] - dy * qx[None, :]) / det
    ok = np.isfinite(t) & (t > 0) & (s >= -1e-9) & (s <= 1 + 1e-9)
    return np.where(ok, t, np.inf).min(axis=1)


@pytest.mark.parametrize("spin,incl", [(0.0, 60.0), (0.5, 17.0), (0.99, 85.0), (0.999, 89.0)])
def test_polygon_radii_star_search_equals_every_segment(spin, incl) -> None:
    """The search over the segment a ray points into (the Bardeen curve is star-shaped about its centroid)
    returns the brute-force minimum bit for bit, rays through vertices and both orientations included."""
    x, y = shadow_curve(Spacetime(1.0, spin), incl, 3000)
    c = polygon_centroid(x, y)
    rng = np.random.default_rng(3)
    ang = np.concatenate([2 * np.pi * np.arange(720) / 720, rng.uniform(-7, 7, 300),
                          np.arctan2(y - c[1], x - c[0])[::29]])
    assert np.array_equal(polygon_radii(x, y, c, ang), _brute_radii(x, y, c, ang))
    assert np.array_equal(polygon_radii(x[::-1], y[::-1], c, ang), _brute_radii(x[::-1], y[::-1], c, ang))


def test_polygon_radii_of_a_non_star_polygon_takes_the_nearest_crossing() -> None:
    """A polygon that folds back (not star-shaped about the centre): every segment is searched."""
    th = np.linspace(0, 2 * np.pi, 400, endpoint=False)
    r = 2.0 + 1.5 * np.sin(5 * th)
    x, y = r * np.cos(th + 0.6 * np.sin(3 * th)), r * np.sin(th + 0.6 * np.sin(3 * th))
    ang = np.linspace(0, 2 * np.pi, 97)
    assert np.array_equal(polygon_radii(x, y, (0.0, 0.0), ang), _brute_radii(x, y, (0.0, 0.0), ang))


def test_oracle_mask_boundary_of_a_synthetic_disk(oracle) -> None:
    cam = BardeenCamera(radius=1000.0, inclination_deg=60.0, fov=10.0, resolution=128)
    alpha, beta = cam.pixel_coordinates()
    captured = (alpha - 0.5) ** 2 + beta**2 < 4.0**2
    ang, rad = oracle.mask_edge(alpha, beta, captured, 90, (0.5, 0.0))
    assert np.max(np.abs(rad - 4.0)) < 0.5 * cam.pixel_size


def test_rounds_for_width() -> None:
    assert rounds_for_width(0.2, 0.2 / 2**34, 1) == 34
    assert rounds_for_width(0.2, 0.2 / 64**5, 63) == 5
