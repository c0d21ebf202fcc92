"""ZAMO camera on the GPU: device initial states, traces and the planned and batched forms.

* ``kr_image_states_dev`` builds every pixel's photon from the camera's
  frame with IEEE multiplies, divisions, square roots and fused
  multiply-adds, no transcendental: bit for bit the reference's
  ``initial_states`` (tests/golden/zamo_cameras.npz) on all ten cameras,
  reversed and physical.
* The 64 x 64 images of docs/shadow.md section 5 (rk45, rtol 1e-9, atol
  1e-11, i = 60 deg, fov 12): 558 / 3538 captured / escaped at a = 0.9,
  608 / 3488 at a = 0, every pixel's state the reference's, and the rows
  within the class-aware tolerance of the oracle (the device's sin/cos/pow
  round differently from glibc's in a few percent of calls).
* The planned forms (FP64, and FP32 with FP64 repair) of a 512 x 512 ZAMO
  image give the FP64 mask of the plain trace; the ragged batch traces ZAMO
  images as the single-image entry point does.
"""

from __future__ import annotations

import os

import numpy as np
import pytest
from conftest import GOLDEN, options
from test_gpu_parity import report, row_differences

from kerrray_amd.camera import Camera, initial_states
from kerrray_amd.geometry import Spacetime

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gold() -> dict:
    with np.load(os.path.join(GOLDEN, "zamo_cameras.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _cameras(gold):
    for k, (spin, incl, phi, fov, n, r_o) in enumerate(gold["cameras"]):
        yield k, Spacetime(1.0, float(spin)), Camera(radius=r_o, inclination_deg=incl, phi_deg=phi, fov=fov,
                                                      resolution=int(n))


def test_device_states_equal_the_reference(hip, gold) -> None:
    for k, st, cam in _cameras(gold):
        print(f"[zamo states] camera {k}: a={st.spin} i={cam.inclination_deg} r_o={cam.radius} n={cam.resolution}",
              flush=True)
        assert np.array_equal(initial_states(cam, st, backend=hip), gold[f"c{k}_Y0"]), k
        assert np.array_equal(initial_states(cam, st, reverse=False, backend=hip), gold[f"c{k}_Y0_physical"]), k


def _numpy_rows() -> dict:
    with np.load(os.path.join(GOLDEN, "zamo64_numpy_rows.npz"), allow_pickle=False) as z:
        return {key: z[key] for key in z.files}


@pytest.mark.parametrize("k, captured, escaped", [(0, 558, 3538), (1, 608, 3488)])
def test_documented_64_images(hip, oracle, gold, k, captured, escaped) -> None:
    from kerrray_amd.shadow import compute_shadow

    _, st, cam = list(_cameras(gold))[k]
    integ, term = options(rtol=1e-9, atol=1e-11)
    img = compute_shadow(st, cam, integ, term, backend=hip)
    counts = (img.counts["CAPTURED"], img.counts["ESCAPED"])
    print(f"[zamo 64^2 a={st.spin}] captured/escaped {counts}, mean steps {img.result.n_steps.mean():.1f}", flush=True)
    assert counts == (captured, escaped) and img.n_failed == 0
    assert np.array_equal(img.result.state, gold[f"c{k}_numpy_state"])
    ref = oracle.integrate_batch(st, gold[f"c{k}_Y0"], integ, term)
    assert np.array_equal(img.result.state, ref.state)
    # At rtol 1e-9 the integration error is below what a one-ulp difference
    # of sin/cos grows to along a photon that circles the hole a few times (a
    # factor of about e^pi per half orbit), so the bound on the final rows is
    # measured, not assumed: the reference's own NumPy integrator differs from
    # its Numba kernel (the oracle's rows, bit for bit) on these images by up
    # to 2.8e-7 (a = 0.9) and 2.2e-7 (a = 0), with the same step sequence
    # (tests/golden/zamo64_numpy_rows.npz).  The device may differ from the
    # oracle by at most twice that, at the maximum, the 99th percentile and
    # the median.
    np_rows = _numpy_rows()
    np_same = (np_rows[f"c{k}_numpy_n_steps"] == ref.n_steps) & (np_rows[f"c{k}_numpy_n_rejected"] == ref.n_rejected)
    spread, _, _, _, _ = row_differences(np_rows[f"c{k}_numpy_Y"], ref.Y, ref.state, spin=st.spin)
    spread = spread[np_same]
    bound = 2.0 * spread.max()
    same, _ = report(f"zamo 64 a={st.spin}", img.result, ref.Y, ref.n_steps, ref.n_rejected, bound / 10.0,
                     ref.state, spin=st.spin)
    mine, _, _, _, _ = row_differences(img.result.Y, ref.Y, ref.state, spin=st.spin)
    mine = mine[same]
    print(f"  reference NumPy vs Numba on the same rows: max {spread.max():.3e} p99 {np.quantile(spread, 0.99):.3e} "
          f"median {np.median(spread):.3e}; device vs oracle: max {mine.max():.3e} p99 {np.quantile(mine, 0.99):.3e} "
          f"median {np.median(mine):.3e}", flush=True)
    assert np.quantile(mine, 0.99) <= 2.0 * np.quantile(spread, 0.99)
    assert np.median(mine) <= 2.0 * np.median(spread)


@pytest.mark.parametrize("fp32_bulk", [False, True])
def test_planned_zamo_image_has_the_fp64_mask(hip, fp32_bulk) -> None:
    from kerrray_amd.planned import PlannedOptions, compute_shadow_planned
    from kerrray_amd.shadow import compute_shadow

    st = Spacetime(1.0, 0.9)
    cam = Camera(radius=1000.0, inclination_deg=60.0, fov=10.0, resolution=512)
    integ, term = options(rtol=1e-6, atol=1e-8)
    plain = compute_shadow(st, cam, integ, term, backend=hip, schedule="plain")
    planned, stats = compute_shadow_planned(st, cam, integ, term, backend=hip,
                                            options=PlannedOptions(fp32_bulk=fp32_bulk))
    diff = int(np.count_nonzero(planned.captured != plain.captured))
    print(f"[zamo planned 512^2 fp32_bulk={fp32_bulk}] mask diff {diff}, repaired {stats.repaired}, "
          f"band {stats.planned_band}, {stats.kernel_ms:.2f} ms", flush=True)
    assert diff == 0
    if not fp32_bulk:
        assert np.array_equal(planned.result.Y, plain.result.Y)
        assert np.array_equal(planned.result.n_steps, plain.result.n_steps)


def test_batch_traces_zamo_images_as_the_single_image_path(hip) -> None:
    from kerrray_amd.batch import DeviceBatch, ImageJob
    from kerrray_amd.shadow import compute_shadow

    jobs = [ImageJob(spin=0.9, inclination_deg=60.0, resolution=64, rtol=1e-7, atol=1e-9, fov=12.0, camera="zamo"),
            ImageJob(spin=0.3, inclination_deg=17.0, resolution=128, rtol=1e-6, atol=1e-8, camera="zamo",
                     phi_deg=20.0),
            ImageJob(spin=0.5, inclination_deg=85.0, resolution=64, rtol=1e-6, atol=1e-8)]
    batch = DeviceBatch(hip, jobs)
    batch.run(graph=False)
    results = batch.results()
    batch.close()
    for job, res in zip(jobs, results):
        st, integ, term = job.options()
        single = compute_shadow(st, job.make_camera(), integ, term, backend=hip).result
        print(f"[zamo batch] {job.camera} a={job.spin} n={job.resolution}: identical "
              f"{np.array_equal(res.Y, single.Y)}", flush=True)
        assert np.array_equal(res.state, single.state) and np.array_equal(res.Y, single.Y)
        assert np.array_equal(res.n_steps, single.n_steps)


def test_points_off_the_sky_raise(hip) -> None:
    import ctypes

    import torch

    from kerrray_amd import _lib
    from kerrray_amd.backends import kernel_params
    from kerrray_amd.camera import image_struct

    st = Spacetime(1.0, 0.9)
    cam = Camera(radius=10.0, inclination_deg=60.0, fov=5.0, resolution=2)
    img = image_struct(st, cam)
    dev = torch.device("cuda:0")
    alpha = torch.tensor([1.0, 9.0], dtype=torch.float64, device=dev)
    beta = torch.tensor([1.0, 9.0], dtype=torch.float64, device=dev)
    integ, term = options(rtol=1e-6, atol=1e-8, lambda_max=50.0)
    p = kernel_params(st, integ, term)
    bufs = {k: torch.zeros(2 * (8 if k == "Y" else 1), dtype=torch.int64 if k in ("n_steps", "n_rejected", "state")
                           else torch.float64, device=dev)
            for k in ("Y", "lam", "n_steps", "n_rejected", "state", "max_null_error", "max_energy_drift",
                      "max_lz_drift", "max_carter_drift")}
    out = _lib.KrOutputs(**{k: v.data_ptr() for k, v in bufs.items()})
    flags = torch.zeros(1, dtype=torch.int32, device=dev)
    _lib.check(_lib.lib().kr_trace_points_dev(hip.handle.ptr, ctypes.byref(p), ctypes.byref(img), alpha.data_ptr(),
                                              beta.data_ptr(), 2, ctypes.byref(out), flags.data_ptr(), None),
               "kr_trace_points_dev")
    torch.cuda.synchronize()
    assert int(flags.item()) & _lib.KR_FLAG_SKY
    with pytest.raises(ValueError):
        _lib.raise_for_flags(int(flags.item()))
    state = bufs["state"].cpu().numpy()
    assert state[1] == 4 and state[0] in (1, 2, 3)  # the point off the sky has no finite photon: FAILURE
