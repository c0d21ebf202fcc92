"""ZAMO camera: initial states of the reference's observer camera (camera.py:225-381).

What pins what:

* The oracle's restatement (``kro_zamo_frame``/``kro_zamo_states``: full
  4 x 4 matrices, BLAS-order fused accumulation) equals the reference's
  ``zamo_tetrad``, ``metric``, ``inverse_metric`` and ``initial_states`` bit
  for bit on ten cameras (tests/golden/zamo_cameras.npz, written by
  tests/golden/make_golden.py from the reference itself). The same products
  evaluated with separate roundings differ from it, so the comparison sees
  the accumulation order.
* The oracle also passes the reference's closed-form checks
  (tests/test_camera.py:168-399): orthonormal tetrad, Schwarzschild limit,
  null states, reversal, co-frame closed form, the Schwarzschild impact
  parameter with the finite-distance factor, Bardeen's far limit and the
  mirror symmetries.
* The library's host code for the frame (``kr_zamo_frame``) and the kernel
  source built for the CPU (tests/hostref) equal the oracle bit for bit; the
  kernel source then traces the two 64 x 64 images of docs/shadow.md
  section 5 to the documented 558 / 3538 and 608 / 3488 captured / escaped
  pixels and to the reference's own per-pixel states.
"""

from __future__ import annotations

import ctypes
import logging
import math
import os
import subprocess

import numpy as np
import pytest
from conftest import GOLDEN, ROOT, options

from kerrray_amd import _lib
from kerrray_amd.backends import HostBatch, kernel_params
from kerrray_amd.camera import Camera, image_struct, zamo_tetrad
from kerrray_amd.geometry import Spacetime

HOSTREF = os.path.join(ROOT, "tests", "hostref")
ETA = np.diag([-1.0, 1.0, 1.0, 1.0])


@pytest.fixture(scope="module")
def gold() -> dict:
    with np.load(os.path.join(GOLDEN, "zamo_cameras.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def hostref():
    subprocess.run(["make", "-s", "-C", HOSTREF], check=True)
    L = ctypes.CDLL(os.path.join(HOSTREF, "libkr_hostref.so"))
    L.krh_image_states.restype = ctypes.c_int
    L.krh_image_states.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    L.krh_trace_image.restype = ctypes.c_int
    L.krh_trace_image.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    return L


def cameras(gold):
    for k, (spin, incl, phi, fov, n, r_o) in enumerate(gold["cameras"]):
        cam = Camera(radius=r_o, inclination_deg=incl, phi_deg=phi, fov=fov, resolution=int(n))
        yield k, Spacetime(1.0, float(spin)), cam


def grid(cam: Camera) -> tuple[np.ndarray, np.ndarray]:
    a, b = cam.pixel_coordinates()
    return a.ravel(), b.ravel()


def oracle_states(oracle, st, cam, reverse=True):
    a, b = grid(cam)
    return oracle.zamo_states(st.mass, st.spin, cam.radius, cam.inclination_rad, cam.phi_rad, a, b, reverse)


# --- the oracle against the reference -------------------------------------------


def test_oracle_equals_the_reference_bit_for_bit(oracle, gold) -> None:
    for k, st, cam in cameras(gold):
        assert cam.inclination_rad == float(gold[f"c{k}_theta_o"])
        F = oracle.zamo_frame(st.mass, st.spin, cam.radius, cam.inclination_rad)
        assert np.array_equal(F.tetrad, gold[f"c{k}_tetrad"])
        assert np.array_equal(F.metric, gold[f"c{k}_metric"])
        assert np.array_equal(F.inverse_metric, gold[f"c{k}_inverse_metric"])
        for reverse, key in ((True, "Y0"), (False, "Y0_physical")):
            Y, bad = oracle_states(oracle, st, cam, reverse)
            assert bad == 0
            assert np.array_equal(Y, gold[f"c{k}_{key}"]), (k, reverse)


def test_separately_rounded_products_miss_the_reference(gold) -> None:
    """p^phi and p_phi with a rounded product before the sum: not the reference's bits."""
    k, st, cam = next(cameras(gold))
    e, g = gold["c0_tetrad"], gold["c0_metric"]
    a, b = grid(cam)
    u_ph = e[0, 3] + (-a / cam.radius) * e[3, 3]
    p_ph = -(e[0, 0] * g[0, 3] + u_ph * g[3, 3])
    assert not np.array_equal(p_ph, gold["c0_Y0"][:, 7])


def test_schwarzschild_frame_is_the_static_observer(oracle) -> None:
    for r, th in ((3.0, 1.0), (1000.0, math.pi / 2), (2.1, 3.0)):
        f = 1.0 - 2.0 / r
        expected = np.diag([1.0 / math.sqrt(f), math.sqrt(f), 1.0 / r, 1.0 / (r * math.sin(th))])
        np.testing.assert_allclose(oracle.zamo_frame(1.0, 0.0, r, th).tetrad, expected, rtol=1e-14, atol=0)


@pytest.mark.parametrize("spin", [0.0, 0.5, -0.9, 0.998])
@pytest.mark.parametrize("r_offset, theta", [(0.1, 0.3), (3.0, math.pi / 2), (10.0, 2.7), (1000.0, 1.0)])
def test_oracle_frame_is_orthonormal_and_complete(oracle, spin, r_offset, theta) -> None:
    r_plus = 1.0 + math.sqrt(1.0 - spin * spin)
    r = r_plus + r_offset if r_offset < 1.0 else r_offset
    F = oracle.zamo_frame(1.0, spin, r, theta)
    np.testing.assert_allclose(F.tetrad @ F.metric @ F.tetrad.T, ETA, rtol=0, atol=1e-12)
    np.testing.assert_allclose(F.tetrad.T @ ETA @ F.tetrad, F.inverse_metric, rtol=1e-12,
                               atol=1e-12 * np.abs(F.inverse_metric).max())
    lapse, omega = 1.0 / F.tetrad[0, 0], F.tetrad[0, 3] / F.tetrad[0, 0]
    assert lapse**2 == pytest.approx(-1.0 / F.inverse_metric[0, 0], rel=1e-12)
    assert omega == pytest.approx(F.inverse_metric[0, 3] / F.inverse_metric[0, 0], rel=1e-12, abs=1e-15)
    assert np.sign(omega) == np.sign(spin)


def _constants(spin, y):
    energy, lz, p_th, th = -y[:, 4], y[:, 7], y[:, 6], y[:, 2]
    carter = p_th**2 + np.cos(th) ** 2 * (lz**2 / np.sin(th) ** 2 - spin**2 * energy**2)
    return energy, lz, carter


def _residual(F, y):
    p = y[:, 4:]
    return np.einsum("ij,ni,nj->n", F.inverse_metric, p, p)


@pytest.mark.parametrize("spin", [0.0, 0.9, -0.7])
@pytest.mark.parametrize("inclination", [1e-3, 17.0, 60.0, 90.0, 150.0])
def test_oracle_states_are_null_reversed_and_laid_out(oracle, spin, inclination) -> None:
    st = Spacetime(1.0, spin)
    cam = Camera(radius=1000.0, inclination_deg=inclination, phi_deg=20.0, resolution=8)
    y, bad = oracle_states(oracle, st, cam)
    yp, badp = oracle_states(oracle, st, cam, reverse=False)
    assert bad == badp == 0
    F = oracle.zamo_frame(1.0, spin, cam.radius, cam.inclination_rad)
    assert np.all(y[:, 0] == 0.0) and np.all(y[:, 1] == cam.radius)
    assert np.all(y[:, 2] == cam.inclination_rad) and np.all(y[:, 3] == cam.phi_rad)
    assert np.abs(_residual(F, y)).max() < 1e-12 and np.abs(_residual(F, yp)).max() < 1e-12
    assert np.all(y[:, 4] > 0.0) and np.all(F.inverse_metric[1, 1] * y[:, 5] < 0.0)
    assert np.all(yp[:, 4] < 0.0) and np.all(F.inverse_metric[1, 1] * yp[:, 5] > 0.0)
    np.testing.assert_allclose(y[:, 4:], -yp[:, 4:], rtol=1e-13, atol=0)


@pytest.mark.parametrize("spin", [0.0, 0.9, -0.7])
@pytest.mark.parametrize("inclination", [30.0, 90.0])
def test_oracle_states_match_the_coframe_closed_form(oracle, spin, inclination) -> None:
    st = Spacetime(1.0, spin)
    cam = Camera(radius=1000.0, inclination_deg=inclination, resolution=6)
    y, _ = oracle_states(oracle, st, cam, reverse=False)
    F = oracle.zamo_frame(1.0, spin, cam.radius, cam.inclination_rad)
    a, b = grid(cam)
    lapse, omega = 1.0 / F.tetrad[0, 0], F.tetrad[0, 3] / F.tetrad[0, 0]
    g = F.metric
    l_ph, l_th = -a / cam.radius, b / cam.radius
    l_r = np.sqrt(1.0 - (a**2 + b**2) / cam.radius**2)
    np.testing.assert_allclose(y[:, 4], -lapse - omega * np.sqrt(g[3, 3]) * l_ph, rtol=1e-12)
    np.testing.assert_allclose(y[:, 5], np.sqrt(g[1, 1]) * l_r, rtol=1e-12)
    np.testing.assert_allclose(y[:, 6], np.sqrt(g[2, 2]) * l_th, rtol=1e-12)
    np.testing.assert_allclose(y[:, 7], np.sqrt(g[3, 3]) * l_ph, rtol=1e-12)


@pytest.mark.parametrize("radius, tol", [(1.0e4, 1e-3), (1.0e6, 2e-6)])
def test_oracle_schwarzschild_impact_parameter(oracle, radius, tol) -> None:
    cam = Camera(radius=radius, inclination_deg=60.0, resolution=8)
    y, _ = oracle_states(oracle, Spacetime(1.0, 0.0), cam)
    energy, lz, carter = _constants(0.0, y)
    b = np.sqrt(lz**2 + carter) / np.abs(energy)
    a0, b0 = grid(cam)
    impact = np.sqrt(a0**2 + b0**2)
    assert np.abs(b / impact - 1.0).max() < tol
    np.testing.assert_allclose(b, impact / math.sqrt(1.0 - 2.0 / radius), rtol=1e-12)


@pytest.mark.parametrize("spin", [0.9, -0.5])
def test_oracle_far_limit_is_bardeen(oracle, spin) -> None:
    cam = Camera(radius=1.0e6, inclination_deg=60.0, resolution=8)
    y, _ = oracle_states(oracle, Spacetime(1.0, spin), cam)
    energy, lz, carter = _constants(spin, y)
    xi, eta = lz / energy, carter / energy**2
    a, b = grid(cam)
    th = math.radians(60.0)
    np.testing.assert_allclose(a, -xi / math.sin(th), rtol=1e-5)
    np.testing.assert_allclose(b, y[:, 6] / energy, rtol=1e-5)
    np.testing.assert_allclose(b**2, eta + spin**2 * math.cos(th) ** 2 - xi**2 / math.tan(th) ** 2, rtol=1e-5)


@pytest.mark.parametrize("spin", [0.0, 0.9, -0.7])
def test_oracle_mirror_symmetries(oracle, spin) -> None:
    cam = Camera(radius=1000.0, inclination_deg=30.0, resolution=8)
    y, _ = oracle_states(oracle, Spacetime(1.0, spin), cam)
    y = y.reshape(cam.shape + (8,))
    F = oracle.zamo_frame(1.0, spin, cam.radius, cam.inclination_rad)
    omega = F.tetrad[0, 3] / F.tetrad[0, 0]
    mirrored = y[:, ::-1, :]
    np.testing.assert_allclose(mirrored[..., 7], -y[..., 7], rtol=1e-12)
    np.testing.assert_allclose(mirrored[..., 5], y[..., 5], rtol=1e-14)
    e_zamo = -(y[..., 4] + omega * y[..., 7])
    np.testing.assert_allclose(e_zamo[:, ::-1], e_zamo, rtol=1e-12)
    flipped = y[::-1, :, :]
    np.testing.assert_allclose(flipped[..., 6], -y[..., 6], rtol=1e-12)
    for c in (4, 5, 7):
        np.testing.assert_allclose(flipped[..., c], y[..., c], rtol=1e-13)


def test_oracle_flags_points_off_the_sky(oracle) -> None:
    _, bad = oracle.zamo_states(1.0, 0.9, 10.0, 1.0, 0.0, [8.0, 1.0], [8.0, 1.0])
    assert bad & 4


# --- the product's host code and the kernel source on the CPU --------------------


def test_library_frame_equals_the_reference(gold) -> None:
    for k, st, cam in cameras(gold):
        np.testing.assert_array_equal(zamo_tetrad(st, cam.radius, cam.inclination_rad), gold[f"c{k}_tetrad"])
        e, g, gi = np.empty(16), np.empty(16), np.empty(16)
        dp = ctypes.POINTER(ctypes.c_double)
        img = image_struct(st, cam)
        _lib.check(_lib.lib().kr_zamo_frame(ctypes.byref(img), e.ctypes.data_as(dp), g.ctypes.data_as(dp),
                                            gi.ctypes.data_as(dp)), "kr_zamo_frame")
        assert np.array_equal(g.reshape(4, 4), gold[f"c{k}_metric"])
        assert np.array_equal(gi.reshape(4, 4), gold[f"c{k}_inverse_metric"])


def test_library_frame_rejects_the_axis_and_the_horizon() -> None:
    st = Spacetime(1.0, 0.9)
    for r, th in ((1.2, 1.0), (3.0, 0.0), (3.0, math.pi), (math.nan, 1.0)):
        with pytest.raises(ValueError):
            zamo_tetrad(st, r, th)


def test_kernel_source_states_equal_the_reference(hostref, gold) -> None:
    for k, st, cam in cameras(gold):
        img = image_struct(st, cam)
        for reverse, key in ((1, "Y0"), (0, "Y0_physical")):
            Y = np.empty((cam.n_rays, 8))
            assert hostref.krh_image_states(ctypes.addressof(img), reverse, Y.ctypes.data) == 0
            assert np.array_equal(Y, gold[f"c{k}_{key}"]), (k, reverse)


def test_kernel_source_rejects_a_field_of_view_beyond_the_sky(hostref) -> None:
    img = image_struct(Spacetime(1.0, 0.9), Camera(radius=10.0, inclination_deg=60.0, fov=9.0, resolution=8))
    Y = np.empty((64, 8))
    assert hostref.krh_image_states(ctypes.addressof(img), 1, Y.ctypes.data) == -1


@pytest.mark.parametrize("k, captured, escaped, mean_steps", [(0, 558, 3538, 234), (1, 608, 3488, 224)])
def test_kernel_source_traces_the_documented_64_images(hostref, oracle, gold, k, captured, escaped,
                                                       mean_steps) -> None:
    """docs/shadow.md section 5: rk45, rtol 1e-9, atol 1e-11, r_o 1000, i 60, fov 12, 64 x 64."""
    _, st, cam = list(cameras(gold))[k]
    integ, term = options(rtol=1e-9, atol=1e-11)
    p = kernel_params(st, integ, term)
    img = image_struct(st, cam)
    out = HostBatch(cam.n_rays, integ.np_dtype, False)
    assert hostref.krh_trace_image(ctypes.addressof(p), ctypes.addressof(img), ctypes.addressof(out.outputs), 0) == 0
    res = out.result(0.0)
    assert (int((res.state == 2).sum()), int((res.state == 1).sum())) == (captured, escaped)
    assert round(float(res.n_steps.mean())) == mean_steps
    assert np.array_equal(res.state, gold[f"c{k}_numpy_state"])
    # The reference's NumPy integrator raises to the -1/5 power with NumPy's
    # own pow, not glibc's: a handful of pixels take one step more or less.
    assert np.mean(res.n_steps == gold[f"c{k}_numpy_n_steps"]) > 0.995
    ref = oracle.integrate_batch(st, gold[f"c{k}_Y0"], integ, term)
    for name in ("Y", "state", "n_steps", "n_rejected", "lam", "max_null_error", "max_carter_drift"):
        assert np.array_equal(getattr(res, name), getattr(ref, name)), name


# --- the Camera dataclass ----------------------------------------------------------


def test_camera_defaults_validation_and_clamp(caplog) -> None:
    cam = Camera(radius=1000, inclination_deg=60)
    assert (cam.phi_deg, cam.fov, cam.resolution, cam.n_rays, cam.shape) == (0.0, 12.0, 64, 4096, (64, 64))
    assert cam.inclination_rad == math.radians(60.0)
    for bad in ({"radius": 0.0}, {"inclination_deg": 180.5}, {"phi_deg": math.inf}, {"fov": -3.0},
                {"resolution": 63}, {"resolution": True}):
        with pytest.raises(ValueError):
            Camera(**{"radius": 1000.0, "inclination_deg": 60.0, **bad})
    with caplog.at_level(logging.WARNING, logger="kerrray_amd"):
        assert Camera(radius=1000.0, inclination_deg=0.0, resolution=4).inclination_deg == 1e-3
        assert Camera(radius=1000.0, inclination_deg=180.0, resolution=4).inclination_deg == 180.0 - 1e-3
    assert sum("clamped" in r.getMessage() for r in caplog.records) == 2
    img = image_struct(Spacetime(1.0, 0.9), Camera(radius=1000.0, inclination_deg=60.0, phi_deg=20.0))
    assert img.camera == _lib.KR_CAMERA_ZAMO and img.phi_o == math.radians(20.0)
