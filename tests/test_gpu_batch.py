"""The persistent, work-stealing batch kernel gives each image exactly the plain kernel's results."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("Y", "state", "n_steps", "n_rejected", "lam", "max_null_error", "max_energy_drift", "max_lz_drift",
          "max_carter_drift")


def _small_ragged():
    from kerrray_amd.batch import ImageJob

    return [
        ImageJob(0.9, 60.0, 48, 1e-6, 1e-8),
        ImageJob(0.0, 17.0, 16, 1e-9, 1e-11),
        ImageJob(0.99, 85.0, 32, 1e-5, 1e-7),
        ImageJob(0.5, 60.0, 64, 1e-10, 1e-12),
        ImageJob(0.3, 85.0, 8, 1e-7, 1e-9),
        ImageJob(0.7, 17.0, 40, 1e-8, 1e-10),
    ]


def _plain(hip, job):
    from kerrray_amd.camera import BardeenCamera
    from kerrray_amd.shadow import compute_shadow

    st, integ, term = job.options()
    cam = BardeenCamera(radius=job.r_o, inclination_deg=job.inclination_deg, fov=job.fov, resolution=job.resolution)
    return compute_shadow(st, cam, integ, term, backend=hip).result


@pytest.mark.parametrize("tune", [(256, 2, 4, 0), (64, 3, 1, 1), (128, 2, 32, 0), (256, 2, 8, 1)])
def test_ragged_batch_equals_plain_kernel_per_image(hip, tune) -> None:
    """Every image's rows equal the plain kernel's, in job order, whatever the processing order (last field)."""
    from kerrray_amd import _lib
    from kerrray_amd.batch import DeviceBatch

    jobs = _small_ragged()
    batch = DeviceBatch(hip, jobs, _lib.KrTune(*tune))
    batch.run(graph=False)
    res = batch.results()
    batch.close()
    for job, got in zip(jobs, res):
        ref = _plain(hip, job)
        for name in FIELDS:
            assert np.array_equal(getattr(got, name), getattr(ref, name)), (job, name)


def test_graph_replays_are_identical_to_eager(hip) -> None:
    """Several hipGraph replays, with eager work and synchronisation between them, all equal the eager run."""
    import torch

    from kerrray_amd.batch import DeviceBatch

    jobs = _small_ragged()[:4]
    batch = DeviceBatch(hip, jobs)
    batch.run(graph=False)
    eager = batch.results()
    for _ in range(3):
        batch.bufs["state"].fill_(-7)
        batch.bufs["Y"].zero_()
        torch.cuda.synchronize()
        batch.run(graph=True)
        torch.ones(1000, device="cuda").sum().item()
        again = batch.results()
        for a, b in zip(eager, again):
            for name in FIELDS:
                assert np.array_equal(getattr(a, name), getattr(b, name)), name
    batch.close()


def test_single_image_batch_uses_uniform_constants_and_matches(hip, golden2) -> None:
    from kerrray_amd.batch import DeviceBatch, ImageJob

    job = ImageJob(0.9, 60.0, 512, 1e-6, 1e-8)
    batch = DeviceBatch(hip, [job])
    batch.run(graph=True)
    (res,) = batch.results()
    ms = batch.timed(repeats=3)
    batch.close()
    print(f"[persistent single image 512^2] {ms:.3f} ms", flush=True)
    assert np.array_equal(res.state, golden2["numpy_state"].astype(np.int64))
    ref = _plain(hip, job)
    for name in FIELDS:
        assert np.array_equal(getattr(res, name), getattr(ref, name)), name


def test_autotune_picks_a_supported_shape(hip) -> None:
    from kerrray_amd.batch import autotune

    tune, table = autotune(hip, _small_ragged()[:3], block_sizes=(64, 256), waves=(2,), intervals=(1, 8), cache=False)
    assert len(table) == 4 and tune.block_size in (64, 256) and tune.refill_interval in (1, 8)
    assert min(r["ms"] for r in table) == next(
        r["ms"] for r in table if (r["block_size"], r["refill_interval"]) == (tune.block_size, tune.refill_interval))


def test_precision_batch_keeps_every_image_fp64_mask(hip) -> None:
    """The precision form on the batch (FP32 bulk, FP64 band, verify/repair): each image's mask equals the
    FP64 batch's; eager runs and graph replays give the same rows."""
    import torch

    from kerrray_amd import _lib
    from kerrray_amd.batch import DeviceBatch
    from kerrray_amd.planned import PlannedOptions

    jobs = _small_ragged()
    exact = DeviceBatch(hip, jobs, _lib.KrTune(256, 2, 8, 1))
    exact.run(graph=False)
    ref = exact.results()
    exact.close()
    for tune in ((256, 2, 8, 1), (64, 3, 4, 0)):
        mixed = DeviceBatch(hip, jobs, _lib.KrTune(*tune), precision=PlannedOptions())
        mixed.run(graph=False)
        got = mixed.results()
        st = mixed.stats()
        print(f"[precision batch {tune}] band {st.planned_band}, repaired {st.repaired} in {st.repair_rounds} "
              f"rounds, fp64 attempts {st.fp64_attempts}", flush=True)
        for job, a, b in zip(jobs, got, ref):
            assert np.array_equal(a.state == 2, b.state == 2), job
            assert np.array_equal(a.state, b.state), job
        for _ in range(2):
            mixed.bufs["state"].fill_(-7)
            torch.cuda.synchronize()
            mixed.run(graph=True)
            again = mixed.results()
            for a, b in zip(got, again):
                for name in FIELDS:
                    assert np.array_equal(getattr(a, name), getattr(b, name)), name
        mixed.close()


def _many_small(n_images: int = 200):
    """More images than the launch keeps in LDS (128): the job table is read from device memory."""
    from kerrray_amd.batch import ImageJob

    rng = np.random.Generator(np.random.PCG64(7))
    return [
        ImageJob(float(rng.uniform(0.0, 0.99)), float(rng.choice([17.0, 60.0, 85.0])), int(rng.choice([8, 12, 16, 24])),
                 10.0 ** -int(e), 1e-2 * 10.0 ** -int(e))
        for e in rng.integers(5, 9, n_images)
    ]


@pytest.mark.parametrize("tune", [(128, 2, 8, 1), (64, 3, 4, 0)])
def test_batch_of_more_than_128_images_equals_plain_kernel_per_image(hip, tune) -> None:
    """200 images in one launch (constants read from the job table in device memory): every image's rows equal
    the plain kernel's."""
    from kerrray_amd import _lib
    from kerrray_amd.batch import DeviceBatch

    jobs = _many_small()
    batch = DeviceBatch(hip, jobs, _lib.KrTune(*tune))
    batch.run(graph=True)
    res = batch.results()
    batch.close()
    for job, got in zip(jobs, res):
        ref = _plain(hip, job)
        for name in FIELDS:
            assert np.array_equal(getattr(got, name), getattr(ref, name)), (job, name)


def test_precision_batch_of_more_than_128_images_keeps_every_fp64_mask(hip) -> None:
    from kerrray_amd import _lib
    from kerrray_amd.batch import DeviceBatch
    from kerrray_amd.planned import PlannedOptions

    jobs = _many_small()
    tune = _lib.KrTune(128, 2, 8, 1)
    exact = DeviceBatch(hip, jobs, tune)
    exact.run(graph=False)
    ref = exact.results()
    exact.close()
    mixed = DeviceBatch(hip, jobs, tune, precision=PlannedOptions())
    mixed.run(graph=True)
    got = mixed.results()
    st = mixed.stats()
    mixed.close()
    print(f"[precision batch, {len(jobs)} images] band {st.planned_band}, repaired {st.repaired} in "
          f"{st.repair_rounds} rounds", flush=True)
    for job, a, b in zip(jobs, got, ref):
        assert np.array_equal(a.state, b.state), job


@pytest.mark.parametrize("jobs_kind", ["all_exact", "one_image"])
def test_precision_batch_edge_cases(hip, jobs_kind) -> None:
    """A precision batch whose every image is below fp32_min_rtol (one exact launch, no FP32 launch beside it),
    and a precision batch of one image (constants in the kernel arguments, both launches): rows bitwise the FP64
    batch's for the first, the FP64 mask for the second, eager and replayed alike."""
    from kerrray_amd import _lib
    from kerrray_amd.batch import DeviceBatch, ImageJob
    from kerrray_amd.planned import PlannedOptions

    if jobs_kind == "all_exact":
        jobs = [ImageJob(0.5, 60.0, 24, 1e-8, 1e-10), ImageJob(0.9, 17.0, 16, 1e-7, 1e-9)]
    else:
        jobs = [ImageJob(0.9, 60.0, 64, 1e-6, 1e-8)]
    tune = _lib.KrTune(128, 2, 8, 1)
    exact = DeviceBatch(hip, jobs, tune)
    exact.run(graph=False)
    ref = exact.results()
    exact.close()
    mixed = DeviceBatch(hip, jobs, tune, precision=PlannedOptions())
    mixed.run(graph=False)
    eager = mixed.results()
    st = mixed.stats()
    mixed.run(graph=True)
    again = mixed.results()
    mixed.close()
    n = sum(j.n_rays for j in jobs)
    if jobs_kind == "all_exact":
        assert st.planned_band == n
    else:
        assert 0 < st.planned_band < n
    for a, b, c in zip(eager, again, ref):
        for name in FIELDS:
            assert np.array_equal(getattr(a, name), getattr(b, name)), name
        assert np.array_equal(a.state, c.state)
        if jobs_kind == "all_exact":
            for name in FIELDS:
                assert np.array_equal(getattr(a, name), getattr(c, name)), name
