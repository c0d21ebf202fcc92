"""Planned traces: the FP64 schedule changes nothing; the FP32 bulk with FP64 repair keeps the FP64 mask."""

from __future__ import annotations

import numpy as np
import pytest
from conftest import options

pytestmark = pytest.mark.gpu

FIELDS = ("Y", "state", "n_steps", "n_rejected", "lam", "max_null_error", "max_energy_drift", "max_lz_drift",
          "max_carter_drift")


def _images(hip, spin, incl, res, rtol, opts):
    from kerrray_amd.camera import BardeenCamera
    from kerrray_amd.geometry import Spacetime
    from kerrray_amd.planned import compute_shadow_planned
    from kerrray_amd.shadow import compute_shadow

    st = Spacetime(1.0, spin)
    integ, term = options(rtol=rtol, atol=1e-2 * rtol)
    cam = BardeenCamera(radius=1000.0, inclination_deg=incl, fov=10.0, resolution=res)
    plain = compute_shadow(st, cam, integ, term, backend=hip, schedule="plain")
    plain_ms = hip.last_kernel_ms
    planned, stats = compute_shadow_planned(st, cam, integ, term, backend=hip, options=opts)
    return plain, plain_ms, planned, stats


@pytest.mark.parametrize("case", [(0.9, 60.0, 512), (0.99, 85.0, 1024)])
def test_fp64_planned_schedule_is_bitwise_the_plain_trace(hip, golden2, case) -> None:
    """512^2 runs the kernel that keeps the DP5 error sums in registers, 1024^2 (over 2^19 pixels) the one
    that parks them in LDS: both are the plain trace bit for bit."""
    from kerrray_amd.planned import PlannedOptions

    spin, incl, res = case
    plain, plain_ms, planned, stats = _images(hip, spin, incl, res, 1e-6, PlannedOptions(fp32_bulk=False))
    print(f"[fp64 planned a={spin} {res}^2] {stats.kernel_ms:.3f} ms (plain {plain_ms:.3f} ms), "
          f"band {stats.planned_band}, repaired {stats.repaired}", flush=True)
    for name in FIELDS:
        assert np.array_equal(getattr(planned.result, name), getattr(plain.result, name)), name
    assert stats.repaired == 0 and stats.fp64_fraction == 1.0
    if res == 512:
        assert np.array_equal(planned.result.state, golden2["numpy_state"].astype(np.int64))


@pytest.mark.parametrize("case", [(0.9, 60.0, 512, 1e-6), (0.99, 85.0, 1024, 1e-6), (0.5, 17.0, 256, 1e-7)])
def test_fp32_bulk_with_fp64_repair_keeps_the_fp64_mask(hip, case) -> None:
    from kerrray_amd.planned import PlannedOptions

    spin, incl, res, rtol = case
    plain, plain_ms, mixed, stats = _images(hip, spin, incl, res, rtol, PlannedOptions(fp32_bulk=True))
    diff = int(np.count_nonzero(mixed.captured != plain.captured))
    esc = (plain.result.state == 1) & (mixed.result.state == 1)
    theta_err = np.abs(mixed.result.Y[esc, 2] - plain.result.Y[esc, 2])
    print(f"[fp32+fp64 a={spin} i={incl} {res}^2 rtol {rtol}] {stats.kernel_ms:.3f} ms (fp64 plain {plain_ms:.3f} ms) "
          f"mask diff {diff}, band {stats.planned_band}, repaired {stats.repaired}, "
          f"fp64 attempts {stats.fp64_fraction:.3f}, escaped theta |err| median {np.median(theta_err):.2e} "
          f"max {theta_err.max():.2e}", flush=True)
    assert diff == 0
    assert np.array_equal(mixed.result.state, plain.result.state)
    assert mixed.n_failed == 0


@pytest.mark.parametrize("case", [(0.9, 60.0, 512, 1e-6), (0.99, 85.0, 1024, 1e-6)])
def test_split_roles_give_the_same_rows(hip, case) -> None:
    """Wavefronts of one precision each (split_roles, with the automatic occupancy: one wavefront per SIMD
    at 512^2, two at 1024^2 with the error sums parked in LDS) trace every photon exactly as mixed
    wavefronts at one and at two per SIMD do, and so do FP64 photons carried by lane pairs (TrigPair:
    the even lane's sincos(theta) and the odd lane's sincos(2 theta) exchanged by DPP) and split roles
    without pairs: a photon's arithmetic depends neither on its neighbours nor on where its values are
    kept."""
    from kerrray_amd.camera import BardeenCamera
    from kerrray_amd.geometry import Spacetime
    from kerrray_amd.planned import PlannedOptions, compute_shadow_planned

    spin, incl, res, rtol = case
    st = Spacetime(1.0, spin)
    integ, term = options(rtol=rtol, atol=1e-2 * rtol)
    cam = BardeenCamera(radius=1000.0, inclination_deg=incl, fov=10.0, resolution=res)
    mixed, s0 = compute_shadow_planned(st, cam, integ, term, backend=hip,
                                       options=PlannedOptions(fp32_bulk=True, split_roles=False, waves_per_simd=2))
    split, s1 = compute_shadow_planned(st, cam, integ, term, backend=hip,
                                       options=PlannedOptions(fp32_bulk=True, split_roles=True))
    one, s2 = compute_shadow_planned(st, cam, integ, term, backend=hip,
                                     options=PlannedOptions(fp32_bulk=True, split_roles=False, waves_per_simd=1))
    # Paired FP64 lanes (automatic at one wavefront per SIMD, forced at two) and never-paired split roles.
    paired, s3 = compute_shadow_planned(st, cam, integ, term, backend=hip,
                                        options=PlannedOptions(fp32_bulk=True, split_roles=2))
    unpaired, s4 = compute_shadow_planned(st, cam, integ, term, backend=hip,
                                          options=PlannedOptions(fp32_bulk=True, split_roles=3))
    print(f"[split roles a={spin} {res}^2] {s1.kernel_ms:.3f} ms (mixed wavefronts {s0.kernel_ms:.3f} ms at two "
          f"per SIMD, {s2.kernel_ms:.3f} ms at one; paired {s3.kernel_ms:.3f} ms, never paired {s4.kernel_ms:.3f} ms), "
          f"repaired {s1.repaired}", flush=True)
    for other, so in ((mixed, s0), (one, s2), (paired, s3), (unpaired, s4)):
        for name in FIELDS:
            assert np.array_equal(getattr(split.result, name), getattr(other.result, name)), name
        assert s1.fp64_attempts == so.fp64_attempts and s1.repaired == so.repaired


@pytest.mark.parametrize("case", [(0.0, 60.0, 256, 1e-8, "bardeen"), (0.5, 17.0, 362, 1e-7, "bardeen"),
                                  (0.99, 85.0, 512, 1e-6, "bardeen"), (0.9, 60.0, 384, 1e-6, "zamo")])
def test_lane_pairs_are_bitwise_single_lanes(hip, case) -> None:
    """FP64 band photons carried by lane pairs (split_roles 2) and by one lane each (split_roles 3) give the
    same rows bit for bit over spins 0 to 0.99, inclinations 17 to 85 deg, rtol 1e-6 to 1e-8 and both
    cameras; the bytes are compared, so signed zeros count too."""
    from kerrray_amd.camera import BardeenCamera, Camera
    from kerrray_amd.geometry import Spacetime
    from kerrray_amd.planned import PlannedOptions, compute_shadow_planned

    spin, incl, res, rtol, kind = case
    st = Spacetime(1.0, spin)
    integ, term = options(rtol=rtol, atol=1e-2 * rtol)
    cam = (Camera(radius=1000.0, inclination_deg=incl, fov=10.0, resolution=res) if kind == "zamo" else
           BardeenCamera(radius=1000.0, inclination_deg=incl, fov=10.0, resolution=res))
    paired, sp = compute_shadow_planned(st, cam, integ, term, backend=hip,
                                        options=PlannedOptions(fp32_bulk=True, split_roles=2))
    single, ss = compute_shadow_planned(st, cam, integ, term, backend=hip,
                                        options=PlannedOptions(fp32_bulk=True, split_roles=3))
    print(f"[pairs a={spin} i={incl} {res}^2 rtol {rtol} {kind}] paired {sp.kernel_ms:.3f} ms, one lane each "
          f"{ss.kernel_ms:.3f} ms, band {sp.planned_band}", flush=True)
    for name in FIELDS:
        a, b = getattr(paired.result, name), getattr(single.result, name)
        assert a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8)), name
    assert sp.fp64_attempts == ss.fp64_attempts and sp.repaired == ss.repaired


def test_grid_fraction_gives_the_same_rows(hip) -> None:
    """The persistent trace kernel launched on 7/8 (automatic at 512^2), all and 3/10 of the wavefront slots
    traces every photon alike (work is drawn from a queue, so no photon depends on the grid); fractions
    outside (0, 1] are rejected."""
    from kerrray_amd.camera import BardeenCamera
    from kerrray_amd.geometry import Spacetime
    from kerrray_amd.planned import PlannedOptions, compute_shadow_planned

    st = Spacetime(1.0, 0.9)
    integ, term = options(rtol=1e-6, atol=1e-8)
    cam = BardeenCamera(radius=1000.0, inclination_deg=60.0, fov=10.0, resolution=512)
    runs = {}
    for frac in (0.0, 1.0, 0.3):
        img, stats = compute_shadow_planned(st, cam, integ, term, backend=hip,
                                            options=PlannedOptions(fp32_bulk=True, grid_fraction=frac))
        runs[frac] = (img, stats)
        print(f"[grid fraction {frac or 'auto'}] {stats.kernel_ms:.3f} ms, repaired {stats.repaired}", flush=True)
    base, s0 = runs[0.0]
    for frac in (1.0, 0.3):
        other, so = runs[frac]
        for name in FIELDS:
            assert np.array_equal(getattr(base.result, name), getattr(other.result, name)), (frac, name)
        assert s0.fp64_attempts == so.fp64_attempts and s0.repaired == so.repaired
    for bad in (-0.1, 1.5, float("nan")):
        with pytest.raises(ValueError, match="grid_fraction"):
            compute_shadow_planned(st, BardeenCamera(radius=1000.0, inclination_deg=60.0, fov=10.0, resolution=64),
                                   integ, term, backend=hip, options=PlannedOptions(grid_fraction=bad))


def test_precision_form_mask_equals_the_oracle_fp64_mask_1024(hip, oracle) -> None:
    """1024^2, a=0.99, i=85 deg, rtol 1e-6: the FP32+FP64-repair mask, the GPU FP64 mask and the CPU oracle's."""
    from kerrray_amd.camera import BardeenCamera
    from kerrray_amd.geometry import Spacetime
    from kerrray_amd.planned import PlannedOptions, compute_shadow_planned

    st = Spacetime(1.0, 0.99)
    integ, term = options(rtol=1e-6, atol=1e-8)
    cam = BardeenCamera(radius=1000.0, inclination_deg=85.0, fov=10.0, resolution=1024)
    mixed, stats = compute_shadow_planned(st, cam, integ, term, backend=hip, options=PlannedOptions(fp32_bulk=True))
    Y0 = oracle.bardeen_image(1.0, 0.99, 1000.0, 85.0, 10.0, 1024)
    ref = oracle.integrate_batch(st, Y0, integ, term, dynamic=True)
    ref_captured = (ref.state == 2).reshape(1024, 1024)
    diff = int(np.count_nonzero(mixed.captured != ref_captured))
    print(f"[precision 1024^2 a=0.99 i=85] mask diff vs oracle FP64 {diff}, band {stats.planned_band}, "
          f"repaired {stats.repaired}, fp64 attempts {stats.fp64_fraction:.3f}", flush=True)
    assert diff == 0


@pytest.mark.parametrize("case", [("bardeen", 0.9, 60.0, 256), ("zamo", -0.5, 30.0, 128), ("zamo", 0.0, 60.0, 64),
                                  ("bardeen", 0.998, 89.0, 96)])
def test_compute_shadow_auto_schedule_is_bitwise_the_plain_kernel(hip, case) -> None:
    """compute_shadow's default schedule (the planned FP64 kernel for Dormand-Prince in float64) returns the plain
    kernel's rows bit for bit, for both cameras, both spin signs and near-extremal spin."""
    from kerrray_amd.camera import BardeenCamera, Camera
    from kerrray_amd.geometry import Spacetime
    from kerrray_amd.shadow import compute_shadow

    kind, spin, incl, res = case
    st = Spacetime(1.0, spin)
    integ, term = options(rtol=1e-7, atol=1e-9)
    cam = (BardeenCamera(radius=1000.0, inclination_deg=incl, fov=10.0, resolution=res) if kind == "bardeen"
           else Camera(radius=1000.0, inclination_deg=incl, fov=10.0, resolution=res))
    auto = compute_shadow(st, cam, integ, term, backend=hip)
    auto_ms = hip.last_kernel_ms
    plain = compute_shadow(st, cam, integ, term, backend=hip, schedule="plain")
    print(f"[compute_shadow {kind} a={spin} i={incl} {res}^2] auto {auto_ms:.3f} ms, plain {hip.last_kernel_ms:.3f} ms",
          flush=True)
    for name in FIELDS:
        assert np.array_equal(getattr(auto.result, name), getattr(plain.result, name)), name
    assert np.array_equal(auto.captured, plain.captured)


def test_precision_form_keeps_the_fp64_states_on_random_images(hip) -> None:
    """32 seeded random images across the parameter space (spin up to 0.999, inclinations 5-89 deg, both
    cameras, resolutions 96-256, fields of view 6-14 M, rtol 1e-5 to 1e-7): the FP32-bulk form's termination
    codes equal the FP64 form's on every pixel, and verify/repair settles within its round limit (the host
    entry point raises otherwise)."""
    from kerrray_amd.camera import BardeenCamera, Camera
    from kerrray_amd.geometry import Spacetime
    from kerrray_amd.planned import PlannedOptions, compute_shadow_planned

    rng = np.random.Generator(np.random.PCG64(11))
    repaired = 0
    for k in range(32):
        spin = float(rng.uniform(0.0, 0.999))
        incl = float(rng.uniform(5.0, 89.0))
        res = int(rng.choice([96, 128, 160, 200, 256]))
        fov = float(rng.choice([6.0, 10.0, 14.0]))
        rtol = float(rng.choice([1e-5, 1e-6, 1e-7]))
        kind = Camera if k % 4 == 3 else BardeenCamera
        cam = kind(radius=1000.0, inclination_deg=incl, fov=fov, resolution=res)
        st = Spacetime(1.0, spin)
        integ, term = options(rtol=rtol, atol=1e-2 * rtol)
        print(f"[random image {k}] {kind.__name__} a={spin:.3f} i={incl:.1f} {res}^2 fov {fov} rtol {rtol}", flush=True)
        fp64, _ = compute_shadow_planned(st, cam, integ, term, backend=hip, options=PlannedOptions(fp32_bulk=False))
        mixed, stats = compute_shadow_planned(st, cam, integ, term, backend=hip, options=PlannedOptions())
        assert np.array_equal(mixed.result.state, fp64.result.state), k
        repaired += stats.repaired
    print(f"[random images] 32 images, {repaired} pixels repaired in total", flush=True)


def test_fp64_planned_rows_are_the_plain_rows_on_random_images(hip) -> None:
    """16 seeded random images (both cameras): the longest-first persistent schedule in FP64 gives every row of
    the one-thread-per-pixel kernel bit for bit."""
    from kerrray_amd.camera import BardeenCamera, Camera
    from kerrray_amd.geometry import Spacetime
    from kerrray_amd.planned import PlannedOptions, compute_shadow_planned
    from kerrray_amd.shadow import compute_shadow

    rng = np.random.Generator(np.random.PCG64(12))
    for k in range(16):
        spin = float(rng.uniform(0.0, 0.999))
        incl = float(rng.uniform(5.0, 89.0))
        res = int(rng.choice([64, 96, 128, 200]))
        fov = float(rng.choice([6.0, 10.0, 14.0]))
        rtol = float(rng.choice([1e-5, 1e-6, 1e-7, 1e-8]))
        kind = Camera if k % 4 == 1 else BardeenCamera
        cam = kind(radius=1000.0, inclination_deg=incl, fov=fov, resolution=res)
        st = Spacetime(1.0, spin)
        integ, term = options(rtol=rtol, atol=1e-2 * rtol)
        print(f"[random image {k}] {kind.__name__} a={spin:.3f} i={incl:.1f} {res}^2 fov {fov} rtol {rtol}", flush=True)
        plain = compute_shadow(st, cam, integ, term, backend=hip, schedule="plain")
        planned, _ = compute_shadow_planned(st, cam, integ, term, backend=hip, options=PlannedOptions(fp32_bulk=False))
        for name in FIELDS:
            assert np.array_equal(getattr(planned.result, name), getattr(plain.result, name)), (k, name)
