"""Physics of shadow images traced on the device, checked as the reference checks its own
(tests/test_shadow.py:56-126).

The shadows are computed once per module with the reference's camera (ZAMO observer at r_o = 1000 M,
i = 60 deg, fov 8 M) at 16x16 and 24x24 (a = 0) and 24x24 (a = 0.9), DP5 rtol 1e-7 / atol 1e-9, on the
hip backend. The reference values (Schwarzschild b_c = 3 sqrt(3) M, Bardeen's curve) come from
``kerrray_amd.orbits`` and are only checked against here.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
from conftest import options

pytestmark = pytest.mark.gpu

R_O, INCL, FOV = 1000.0, 60.0, 8.0
IMAGES = {"schwarzschild_16": (0.0, 16), "schwarzschild_24": (0.0, 24), "kerr_24": (0.9, 24)}
SCHWARZSCHILD = ("schwarzschild_16", "schwarzschild_24")


@pytest.fixture(scope="module")
def shadows(hip):
    from kerrray_amd.camera import Camera
    from kerrray_amd.geometry import Spacetime
    from kerrray_amd.shadow import compute_shadow

    integ, term = options(rtol=1e-7, atol=1e-9)
    out = {}
    for name, (spin, res) in IMAGES.items():
        cam = Camera(radius=R_O, inclination_deg=INCL, fov=FOV, resolution=res)
        out[name] = compute_shadow(Spacetime(1.0, spin), cam, integ, term, backend=hip)
    return out


def _boundary(hip, img, centre=None):
    from kerrray_amd.boundary import extract_boundary_from_mask, mask_centroid

    c = mask_centroid(img.alpha, img.beta, img.captured) if centre is None else centre
    angles, radii = extract_boundary_from_mask(img.alpha, img.beta, img.captured, n_angles=360, centre=c,
                                               backend=hip)
    return c, angles, radii


@pytest.mark.parametrize("name", list(IMAGES))
def test_every_pixel_is_captured_or_escaped_and_counts_sum(shadows, name) -> None:
    from kerrray_amd.photons import TerminationState
    from kerrray_amd.shadow import state_counts

    img = shadows[name]
    assert sum(img.counts.values()) == img.n_rays == img.camera.resolution**2
    assert img.count(TerminationState.CAPTURED) + img.count(TerminationState.ESCAPED) == img.n_rays
    assert img.n_failed == 0 and img.n_other == 0
    assert img.count(TerminationState.CAPTURED) > 0
    assert np.array_equal(img.captured, img.state == int(TerminationState.CAPTURED))
    assert img.counts == state_counts(img.state)


@pytest.mark.parametrize("name", SCHWARZSCHILD)
def test_schwarzschild_captured_region_is_a_disk(shadows, name) -> None:
    """Spherical symmetry: capture depends only on the pixel radius."""
    img = shadows[name]
    radius = np.hypot(img.alpha, img.beta)
    assert radius[img.captured].max() < radius[~img.captured].min()


@pytest.mark.parametrize("name", SCHWARZSCHILD)
def test_schwarzschild_radius_matches_critical_impact_parameter(hip, shadows, name) -> None:
    """Mean edge radius = b_c within one pixel plus the O(M/r_o) finite-distance offset; a disk to within a
    pixel."""
    from kerrray_amd.orbits import critical_impact_parameters

    img = shadows[name]
    b_c = critical_impact_parameters(img.spacetime)[0]
    assert b_c == pytest.approx(3.0 * math.sqrt(3.0), rel=1e-10)
    _, angles, radii = _boundary(hip, img)
    assert angles.shape == radii.shape == (360,)
    offset = b_c * (1.0 - math.sqrt(1.0 - 2.0 / R_O))
    assert abs(radii.mean() - b_c) < img.pixel_size + offset
    assert radii.max() - radii.min() < img.pixel_size


def test_kerr_shadow_boundary_matches_bardeen_curve(hip, shadows) -> None:
    """a = 0.9: edge within 1.5 px rms (3 px max) of Bardeen's curve about the mask centroid, displaced to
    alpha > 0 like the curve, centroids within a pixel (the reference's boundary_error definitions)."""
    from kerrray_amd.boundary import finite_distance_scale, polygon_centroid, polygon_radii
    from kerrray_amd.orbits import shadow_curve

    img = shadows["kerr_24"]
    c, angles, radii = _boundary(hip, img)
    curve = shadow_curve(img.spacetime, INCL, 720)
    diff = radii - polygon_radii(curve[0], curve[1], c, angles)
    px = img.pixel_size
    assert math.sqrt(np.mean(diff**2)) / px < 1.5
    assert np.max(np.abs(diff)) / px < 3.0
    ax, _ = polygon_centroid(*curve)
    nx, ny = polygon_centroid(c[0] + radii * np.cos(angles), c[1] + radii * np.sin(angles))
    assert ax > 0.0 and math.copysign(1.0, nx) == math.copysign(1.0, ax)
    assert math.hypot(nx - ax, ny - polygon_centroid(*curve)[1]) / px < 1.0
    assert finite_distance_scale(img.spacetime, R_O) == pytest.approx(1.0 / math.sqrt(1.0 - 2.0 / R_O))


@pytest.mark.parametrize("name", list(IMAGES))
def test_captured_mask_is_symmetric_under_beta_reflection(shadows, name) -> None:
    """beta -> -beta flips p_theta only; E, L_z and Q are unchanged, so capture is invariant."""
    img = shadows[name]
    assert np.array_equal(img.captured[::-1, :], img.captured)
    assert np.array_equal(img.beta[::-1, :], -img.beta)


@pytest.mark.parametrize("name", list(IMAGES))
def test_conserved_momenta_are_exact_and_layout_is_row_major(shadows, name) -> None:
    """p_t and p_phi are exact constants of the Hamiltonian form (their right-hand side is identically 0)."""
    img = shadows[name]
    assert img.result.max_energy_drift.max() == 0.0
    assert img.result.max_lz_drift.max() == 0.0
    assert np.array_equal(img.result.state.reshape(img.shape), img.state)
    escaped = img.result.state == 1
    assert img.result.max_null_error[escaped].max() < 1e-5


def test_save_and_load_round_trip(shadows, tmp_path) -> None:
    """A device-traced image survives the npz format (shadow.py save/load)."""
    from kerrray_amd.shadow import ShadowImage

    img = shadows["schwarzschild_16"]
    path = img.save(tmp_path / "nested" / "shadow.npz")
    loaded = ShadowImage.load(path)
    assert loaded.camera == img.camera
    assert loaded.spacetime == img.spacetime
    assert loaded.counts == img.counts
    for name in ("alpha", "beta", "state", "captured"):
        assert np.array_equal(getattr(loaded, name), getattr(img, name))
    for name in ("Y", "state", "n_steps", "lam", "max_null_error", "n_rejected"):
        assert np.array_equal(getattr(loaded.result, name), getattr(img.result, name))
    assert loaded.result.event_Y is None
