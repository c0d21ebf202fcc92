"""Accretion-disk rendering: the disk event on the image path and the shading of crossings.

What pins what:

* The oracle's shading (``kro_disk_shade``: Keplerian emitter, four-component
  contraction, emissivity and g^n) against the reference's own
  ``render_disk`` output (tests/golden/disk_a09_i75_64.npz, written by
  tests/golden/make_golden.py): within a few ulp. The residue comes from the
  reference's NumPy ``power`` (r^1.5, g^4, (r / r_in)^-p), which differs from
  glibc's ``pow`` by an ulp in about 5 % of calls.
* The oracle against the closed forms of physics/redshift.py:16-35:
  g = 1 / [u^t (1 - Omega L_z / E)], the Schwarzschild limit
  g = sqrt(1 - 3M/r) / (1 - Omega L_z / E), and invariance under the
  reversal of the momentum.
* The oracle's disk-plane event equals the reference's Numba kernel (run
  interpreted, glibc pow) bit for bit on all 4096 rays of the measured run of
  docs/rendering.md section 6: states, step counts and crossing states.
* The kernel source built for the CPU shades exactly as the oracle does (bit
  for bit: both call glibc), and its full render of that run gives the
  documented 3182 / 54 / 860 ESCAPED / CAPTURED / DISK_HIT pixels, the
  Numba kernel's crossings bit for bit, and the documented g and I_obs
  statistics.
"""

from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np
import pytest
from conftest import GOLDEN, ROOT, options

from kerrray_amd import _lib
from kerrray_amd.backends import HostBatch, kernel_params
from kerrray_amd.camera import Camera, image_struct
from kerrray_amd.geometry import Spacetime
from kerrray_amd.orbits import isco_radius
from kerrray_amd.renderer import DiskModel, DiskParams, event_options

HOSTREF = os.path.join(ROOT, "tests", "hostref")


@pytest.fixture(scope="module")
def gold() -> dict:
    with np.load(os.path.join(GOLDEN, "disk_a09_i75_64.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def hostref():
    subprocess.run(["make", "-s", "-C", HOSTREF], check=True)
    L = ctypes.CDLL(os.path.join(HOSTREF, "libkr_hostref.so"))
    L.krh_disk_shade.restype = ctypes.c_int
    L.krh_disk_shade.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_int64,
                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.krh_image_states.restype = ctypes.c_int
    L.krh_image_states.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    L.krh_integrate.restype = ctypes.c_int
    L.krh_integrate.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int]
    return L


def model(gold, **kw) -> DiskModel:
    args = dict(r_in=float(gold["r_in"]), r_out=float(gold["r_out"]), emissivity_index=float(gold["emissivity_index"]))
    args.update(kw)
    return DiskModel(**args)


def host_shade(L, disk: DiskModel, spin: float, y):
    y = np.ascontiguousarray(y, dtype=np.float64)
    n = len(y)
    I, g, r = np.empty(n), np.empty(n), np.empty(n)
    kd = disk.kr()
    bad = L.krh_disk_shade(ctypes.byref(kd), 1.0, spin, y.ctypes.data, n, I.ctypes.data, g.ctypes.data, r.ctypes.data)
    return bad, g, I, r


def rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


# --- the oracle against the reference ---------------------------------------------


def test_oracle_shading_equals_the_reference_to_a_few_ulp(oracle, gold) -> None:
    hit = gold["hit"]
    cases = [
        (0.9, float(gold["r_in"]), 3.0, 4, True, gold["y_hit"], gold["g"][hit], gold["intensity"][hit]),
        (0.9, float(gold["r_in"]), 2.5, 3, False, gold["y_retro"], gold["g_retro"], gold["intensity_retro"]),
    ]
    for spin, r_in, p, n, pro, y, g_ref, i_ref in cases:
        g, I = oracle.disk_shade(1.0, spin, r_in, float(gold["r_out"]), p, n, pro, y)
        assert np.mean(g == g_ref) > 0.95 and rel(g, g_ref).max() < 2e-15
        assert np.mean(I == i_ref) > 0.85 and rel(I, i_ref).max() < 5e-15
    g, _ = oracle.disk_shade(1.0, 0.0, 1.0, 100.0, 3.0, 4, True, gold["y_schw"])
    assert rel(g, gold["g_schw"]).max() < 2e-15


@pytest.mark.parametrize("spin, prograde", [(0.9, True), (0.9, False), (-0.5, True), (0.0, True)])
def test_oracle_redshift_closed_form_and_reversal(oracle, gold, spin, prograde) -> None:
    y = gold["y_hit"]
    y = y[y[:, 1] > 5.0]
    g, _ = oracle.disk_shade(1.0, spin, 1.0, 100.0, 3.0, 4, prograde, y)
    g_rev, _ = oracle.disk_shade(1.0, spin, 1.0, 100.0, 3.0, 4, prograde,
                                 np.concatenate([y[:, :4], -y[:, 4:]], axis=1))
    np.testing.assert_allclose(g_rev, g, rtol=1e-15)
    m, a, r = 1.0, spin, y[:, 1]
    s = 1.0 if prograde else -1.0
    omega = s * (-1.0 if a < 0 else 1.0) * math.sqrt(m) / (r**1.5 + s * abs(a) * math.sqrt(m))
    g_tt = -(1.0 - 2.0 * m / r)
    g_tph = -2.0 * m * a / r
    g_pp = r * r + a * a + 2.0 * m * a * a / r
    u_t = 1.0 / np.sqrt(-(g_tt + 2.0 * omega * g_tph + omega**2 * g_pp))
    E, lz = -y[:, 4], y[:, 7]
    np.testing.assert_allclose(g, 1.0 / (u_t * (1.0 - omega * lz / E)), rtol=1e-12)
    if spin == 0.0:
        np.testing.assert_allclose(g, np.sqrt(1.0 - 3.0 / r) / (1.0 - np.sqrt(1.0 / r**3) * lz / E), rtol=1e-12)


def test_oracle_rejects_what_the_reference_rejects(oracle, gold) -> None:
    y = gold["y_hit"][:4].copy()
    y[0, 2] += 1e-3  # off the equator
    with pytest.raises(ValueError):
        oracle.disk_shade(1.0, 0.9, 1.0, 20.0, 3.0, 4, True, y)
    inner = gold["y_hit"][gold["y_hit"][:, 1] < 3.0][:4]  # inside the retrograde photon orbit
    with pytest.raises(ValueError):
        oracle.disk_shade(1.0, 0.9, 1.0, 20.0, 3.0, 4, False, inner)


# --- host-side model, and the kernel source on the CPU ------------------------------


def test_isco_and_disk_model_defaults(gold) -> None:
    isco = [isco_radius(Spacetime(1.0, a), prograde=p) for a in (0.0, 0.5, 0.9, -0.7) for p in (True, False)]
    assert np.array_equal(np.asarray(isco), gold["isco"])
    disk = DiskModel.from_params(Spacetime(1.0, 0.9), DiskParams())
    assert disk.r_in == float(gold["r_in"]) and round(disk.r_in, 5) == 2.32088
    assert disk.redshift_power == 4 and event_options(disk).r_out == 20.0
    with pytest.raises(ValueError):
        DiskModel.from_params(Spacetime(1.0, 0.9), DiskParams(r_in=1.5))  # inside the photon orbit
    for bad in ({"r_in": 5.0, "r_out": 4.0}, {"r_in": 2.0, "r_out": 20.0, "intensity_law": "g5"}):
        with pytest.raises(ValueError):
            DiskModel(**bad)
    with pytest.raises(ValueError):
        DiskParams(intensity_law="g2")


def test_kernel_source_shades_as_the_oracle(hostref, oracle, gold) -> None:
    for spin, disk, y in ((0.9, model(gold), gold["y_hit"]),
                          (0.9, model(gold, emissivity_index=2.5, intensity_law="g3", prograde=False), gold["y_retro"]),
                          (0.0, DiskModel(r_in=1.0, r_out=100.0), gold["y_schw"])):
        bad, g, I, r = host_shade(hostref, disk, spin, y)
        g_o, I_o = oracle.disk_shade(1.0, spin, disk.r_in, disk.r_out, disk.emissivity_index, disk.redshift_power,
                                     disk.prograde, y)
        assert bad == 0
        assert np.array_equal(g, g_o) and np.array_equal(I, I_o) and np.array_equal(r, y[:, 1])
    y = gold["y_hit"][:2].copy()
    y[1, 2] = 1.0
    assert host_shade(hostref, model(gold), 0.9, y)[0] == _lib.KR_FLAG_DISK


def test_kernel_source_renders_the_measured_run(hostref, oracle, gold) -> None:
    """docs/rendering.md section 6: a = 0.9, i = 75 deg, 64 x 64, fov 25, rtol 1e-9, disk defaults."""
    st = Spacetime(1.0, 0.9)
    cam = Camera(radius=1000.0, inclination_deg=75.0, fov=25.0, resolution=64)
    disk = model(gold)
    img = image_struct(st, cam)
    Y0 = np.empty((cam.n_rays, 8))
    assert hostref.krh_image_states(ctypes.addressof(img), 1, Y0.ctypes.data) == 0
    integ, term = options(rtol=1e-9, atol=1e-11)
    p = kernel_params(st, integ, term, event_options(disk))
    out = HostBatch(cam.n_rays, integ.np_dtype, True)
    assert hostref.krh_integrate(ctypes.addressof(p), Y0.ctypes.data, len(Y0), ctypes.addressof(out.outputs), 0) == 0
    res = out.result(0.0)
    counts = tuple(int((res.state == k).sum()) for k in (1, 2, 6))
    assert counts == (3182, 54, 860)
    assert np.array_equal(res.state, gold["state"])
    hit = res.event_hit.astype(bool)
    assert np.array_equal(hit, gold["hit"])
    # The Numba kernel's trajectories, bit for bit.
    assert np.array_equal(res.state, gold["numba_state"]) and np.array_equal(res.n_steps, gold["numba_n_steps"])
    assert np.array_equal(res.n_rejected, gold["numba_n_rejected"])
    assert np.array_equal(res.event_Y[hit], gold["numba_event_Y_hit"])
    bad, g, I, _ = host_shade(hostref, disk, 0.9, res.event_Y[hit])
    assert bad == 0
    # The NumPy run's step sizes differ by an ulp where its pow does, so its
    # crossings differ from these by the integration error (rtol 1e-9).
    assert rel(g, gold["g"][hit]).max() < 1e-9
    assert (round(g.min(), 6), round(g.max(), 5), round(g.mean(), 5)) == (0.238877, 1.34898, 0.88578)
    assert round(I.max(), 5) == 2.18889
    ref = oracle.integrate_batch(st, Y0, integ, term, event_options(disk))
    assert np.array_equal(res.event_Y, ref.event_Y) and np.array_equal(res.state, ref.state)


def test_oracle_disk_event_equals_the_numba_kernel(oracle, gold) -> None:
    """The oracle's disk-plane event (numba_kernel.py:439-454) on the 4096 ZAMO rays of the measured run."""
    st = Spacetime(1.0, 0.9)
    cam = Camera(radius=1000.0, inclination_deg=75.0, fov=25.0, resolution=64)
    a, b = cam.pixel_coordinates()
    Y0, bad = oracle.zamo_states(1.0, 0.9, cam.radius, cam.inclination_rad, 0.0, a.ravel(), b.ravel())
    assert bad == 0
    integ, term = options(rtol=1e-9, atol=1e-11)
    res = oracle.integrate_batch(st, Y0, integ, term, event_options(model(gold)))
    hit = gold["numba_event_hit"].astype(bool)
    assert np.array_equal(res.state, gold["numba_state"]) and np.array_equal(res.n_steps, gold["numba_n_steps"])
    assert np.array_equal(res.event_hit.astype(bool), hit)
    assert np.array_equal(res.event_Y[hit], gold["numba_event_Y_hit"])
