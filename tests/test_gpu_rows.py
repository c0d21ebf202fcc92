"""Given initial states (the reference's trace_rays -> integrate_batch path) on the longest-first kernel.

Dormand-Prince batches of at least 2^15 rows (FP64) or 2^17 rows (float32)
run on a persistent kernel with a work list ordered from each row's
constants of motion
(kr_mixed.hip: plan_row). The order changes no photon's arithmetic: every
row must equal the plain one-photon-per-lane kernel's bit for bit, whatever
the order of the rows and whichever camera made them.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
from conftest import options

pytestmark = pytest.mark.gpu

FIELDS = ("Y", "state", "n_steps", "n_rejected", "lam", "max_null_error", "max_energy_drift", "max_lz_drift",
          "max_carter_drift")


def _plain_image(hip, st, cam, integ, term):
    from kerrray_amd.shadow import compute_shadow

    img = compute_shadow(st, cam, integ, term, backend=hip, schedule="plain")
    return img.result, hip.last_kernel_ms


@pytest.mark.parametrize("camera,dtype", [("bardeen", "float64"), ("zamo", "float64"), ("bardeen", "float32"),
                                          ("zamo", "float32")])
def test_rows_equal_the_plain_kernel_at_512(hip, golden2, camera, dtype) -> None:
    """FP64 rows run on planned_trace_kernel, float32 rows on rows_refill_kernel (2^18 rows: both planned)."""
    from kerrray_amd.camera import BardeenCamera, Camera, bardeen_initial_states, initial_states
    from kerrray_amd.geometry import Spacetime

    st = Spacetime(1.0, 0.9)
    if dtype == "float64":
        integ, term = options(rtol=1e-6, atol=1e-8)
    else:
        integ, term = options(rtol=1e-6, atol=1e-8, dtype="float32", horizon_epsilon=1e-3)
    if camera == "bardeen":
        cam = BardeenCamera(radius=1000.0, inclination_deg=60.0, fov=10.0, resolution=512)
        Y0 = bardeen_initial_states(cam, st)
    else:
        cam = Camera(radius=1000.0, inclination_deg=60.0, fov=10.0, resolution=512)
        Y0 = initial_states(cam, st)
    ref, plain_ms = _plain_image(hip, st, cam, integ, term)
    rows = hip.integrate_batch(st, Y0, integ, term)
    rows_ms = hip.last_kernel_ms
    perm = np.random.default_rng(7).permutation(len(Y0))
    shuffled = hip.integrate_batch(st, Y0[perm], integ, term)
    print(f"[rows {camera} {dtype} 512^2] longest-first rows {rows_ms:.3f} ms, shuffled {hip.last_kernel_ms:.3f} ms, "
          f"plain image kernel {plain_ms:.3f} ms", flush=True)
    assert rows.Y.dtype == integ.np_dtype
    for name in FIELDS:
        assert np.array_equal(getattr(rows, name), getattr(ref, name)), name
        assert np.array_equal(getattr(shuffled, name), getattr(ref, name)[perm]), name
    if camera == "bardeen" and dtype == "float64":
        assert np.array_equal(rows.state, golden2["numpy_state"].astype(np.int64))


@pytest.mark.parametrize("seed", range(4))
def test_float32_rows_equal_the_plain_kernel_on_random_images(hip, seed) -> None:
    """Seeded spins, inclinations, fields of view and tolerances; float32 rows of both cameras at 2^17+ rows."""
    from kerrray_amd.camera import BardeenCamera, Camera, bardeen_initial_states, initial_states
    from kerrray_amd.geometry import Spacetime

    rng = np.random.default_rng(2000 + seed)
    st = Spacetime(1.0, float(rng.uniform(0.0, 0.99)))
    incl, fov = float(rng.uniform(5.0, 89.0)), float(rng.uniform(8.0, 30.0))
    rtol = float(10.0 ** rng.uniform(-6.0, -4.0))
    integ, term = options(rtol=rtol, atol=1e-2 * rtol, dtype="float32", horizon_epsilon=1e-3)
    if seed % 2 == 0:
        cam = BardeenCamera(radius=1000.0, inclination_deg=incl, fov=fov, resolution=366)
        Y0 = bardeen_initial_states(cam, st)
    else:
        cam = Camera(radius=1000.0, inclination_deg=incl, fov=fov, resolution=364)
        Y0 = initial_states(cam, st, backend=hip)
    assert len(Y0) >= 1 << 17
    ref, plain_ms = _plain_image(hip, st, cam, integ, term)
    rows = hip.integrate_batch(st, Y0, integ, term)
    print(f"[rows float32 random {seed}] a={st.spin:.3f} i={incl:.1f} fov={fov:.1f} rtol={rtol:.1e} "
          f"{type(cam).__name__}: longest-first rows {hip.last_kernel_ms:.3f} ms, plain image {plain_ms:.3f} ms",
          flush=True)
    for name in FIELDS:
        assert np.array_equal(getattr(rows, name), getattr(ref, name)), name


def test_device_rows_entry_point_equals_the_host_one(hip) -> None:
    import torch

    from bench import device_outputs
    from kerrray_amd import _lib
    from kerrray_amd.backends import kernel_params
    from kerrray_amd.camera import BardeenCamera, bardeen_initial_states
    from kerrray_amd.geometry import Spacetime

    st = Spacetime(1.0, 0.5)
    integ, term = options(rtol=1e-7, atol=1e-9)
    cam = BardeenCamera(radius=1000.0, inclination_deg=17.0, fov=10.0, resolution=256)
    Y0 = bardeen_initial_states(cam, st)
    host = hip.integrate_batch(st, Y0, integ, term)
    bufs, outs = device_outputs(torch, len(Y0), torch.float64)
    y0 = torch.from_numpy(Y0).cuda()
    p = kernel_params(st, integ, term)
    for _ in range(2):  # the second call reuses the handle's device workspace
        _lib.check(_lib.lib().kr_integrate_batch_dev(hip.handle.ptr, ctypes.byref(p), y0.data_ptr(), len(Y0),
                                                     ctypes.byref(outs), None), "kr_integrate_batch_dev")
        torch.cuda.synchronize()
        for name, key in (("Y", "Y"), ("state", "state"), ("n_steps", "n_steps"), ("n_rejected", "n_rejected"),
                          ("lam", "lam"), ("max_null_error", "m0"), ("max_carter_drift", "m3")):
            assert np.array_equal(bufs[key].cpu().numpy(), getattr(host, name)), name


def test_image_plane_states_of_pixel_centres_are_the_camera_states(hip, oracle) -> None:
    """kr_point_states_dev at the pixel centres equals kr_image_states_dev bit for bit; at random points it
    equals the oracle's restatement of the camera (zamo_states) bit for bit."""
    from kerrray_amd.camera import Camera, image_plane_states, initial_states
    from kerrray_amd.geometry import Spacetime

    st = Spacetime(1.0, 0.9)
    cam = Camera(radius=1000.0, inclination_deg=60.0, fov=12.0, resolution=32)
    a, b = cam.pixel_coordinates()
    assert np.array_equal(image_plane_states(cam, st, a.ravel(), b.ravel()), initial_states(cam, st))
    pts = np.random.default_rng(0).uniform(-12.0, 12.0, size=(5000, 2))
    got = image_plane_states(cam, st, pts[:, 0], pts[:, 1])
    ref, bad = oracle.zamo_states(1.0, 0.9, 1000.0, cam.inclination_rad, 0.0, pts[:, 0], pts[:, 1])
    assert bad == 0 and np.array_equal(got, ref)
