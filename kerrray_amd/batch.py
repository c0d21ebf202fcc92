"""Ragged batches of images in one persistent, work-stealing launch.

Each image is a Bardeen image with its own spin, inclination, resolution and
tolerance. :class:`DeviceBatch` binds a batch to device output arrays once
(``kr_batch_create``) and then runs it either as a plain launch or as one
hipGraph replay (``kr_batch_graph_launch``: counter reset + persistent kernel
captured once). Results are identical to tracing each image alone with the
plain kernel (``compute_shadow``), bit for bit.

:func:`autotune` times the launch shapes the kernel supports (threads per
block, resident wavefronts per SIMD, attempts between lane refills) on the
actual batch and returns the fastest; results can be kept in a JSON cache.

:func:`ragged_workload` draws the benchmark batch: 64 images, spins
U[0, 0.99], inclinations from {17, 60, 85} deg, resolutions from {64, 128,
256, 512}^2, rtol log-uniform over {1e-5, ..., 1e-10}, atol = 1e-2 rtol,
r_o = 1000 M, alpha, beta in [-10, 10] M, all float64, from NumPy's PCG64
with seed 0.
"""

from __future__ import annotations

import ctypes
import itertools
import json
import math
import os
import time
from dataclasses import asdict, dataclass

import numpy as np

from kerrray_amd import _lib
from kerrray_amd.backends import HipBackend, kernel_params
from kerrray_amd.camera import BardeenCamera, Camera, image_struct
from kerrray_amd.geodesics import IntegratorOptions, TerminationOptions
from kerrray_amd.geometry import Spacetime
from kerrray_amd.photons import BatchResult

__all__ = ["BLOCK_SIZES", "DeviceBatch", "ImageJob", "REFILL_INTERVALS", "WAVES_PER_SIMD", "autotune", "ragged_workload"]

BLOCK_SIZES = (64, 128, 256)
FP32_MIN_RTOL = 1e-6
"""Images of a precision batch with a smaller rtol are traced in FP64: below it the FP32 stages' round-off
holds the error estimate above what the tolerance accepts (single images: 512^2 at rtol 1e-6 gains 1.3x,
256^2 at 1e-7 gains nothing, 1e-8 loses)."""
WAVES_PER_SIMD = (2, 3)
REFILL_INTERVALS = (1, 2, 4, 8, 16, 32)


@dataclass(frozen=True)
class ImageJob:
    spin: float
    inclination_deg: float
    resolution: int
    rtol: float
    atol: float
    fov: float = 10.0
    r_o: float = 1000.0
    mass: float = 1.0
    camera: str = "bardeen"  # "bardeen" (image-plane photons) or "zamo" (the observer camera)
    phi_deg: float = 0.0

    def __post_init__(self) -> None:
        if self.camera not in ("bardeen", "zamo"):
            raise ValueError(f"camera must be 'bardeen' or 'zamo', got {self.camera!r}")

    @property
    def n_rays(self) -> int:
        return self.resolution * self.resolution

    def make_camera(self) -> Camera | BardeenCamera:
        kind = Camera if self.camera == "zamo" else BardeenCamera
        return kind(radius=self.r_o, inclination_deg=self.inclination_deg, phi_deg=self.phi_deg, fov=self.fov,
                    resolution=self.resolution)

    def options(self) -> tuple[Spacetime, IntegratorOptions, TerminationOptions]:
        return (Spacetime(self.mass, self.spin), IntegratorOptions(method="rk45", rtol=self.rtol, atol=self.atol),
                TerminationOptions())


def ragged_workload(seed: int = 0, n_images: int = 64) -> list[ImageJob]:
    rng = np.random.Generator(np.random.PCG64(seed))
    spins = rng.uniform(0.0, 0.99, n_images)
    incl = rng.choice(np.array([17.0, 60.0, 85.0]), n_images)
    res = rng.choice(np.array([64, 128, 256, 512]), n_images)
    expo = rng.integers(5, 11, n_images)  # rtol = 10^-5 ... 10^-10, uniform in the exponent
    return [
        ImageJob(spin=float(s), inclination_deg=float(i), resolution=int(r), rtol=10.0 ** (-int(e)),
                 atol=1e-2 * 10.0 ** (-int(e)))
        for s, i, r, e in zip(spins, incl, res, expo)
    ]


def _kr_job(job: ImageJob) -> _lib.KrJob:
    st, integ, term = job.options()
    k = _lib.KrJob()
    k.params = kernel_params(st, integ, term)
    k.image = image_struct(st, job.make_camera())
    return k


class DeviceBatch:
    """A batch bound to device outputs (torch tensors on ``cuda:<device>``)."""

    def __init__(self, backend: HipBackend, jobs: list[ImageJob], tune: _lib.KrTune | None = None, *,
                 precision=None, fp32_min_rtol: float = FP32_MIN_RTOL) -> None:
        """``precision``: a :class:`kerrray_amd.planned.PlannedOptions` (its ``band_px`` and ``axis_lz``) to run
        the precision form (FP32 bulk, FP64 band, verify/repair to the FP64 mask) on every image whose rtol is
        at least ``fp32_min_rtol`` (the others are traced in FP64); ``None``: exact FP64 throughout."""
        import torch

        self.torch = torch
        self.backend = backend
        self.jobs = list(jobs)
        self.offsets = np.concatenate([[0], np.cumsum([j.n_rays for j in self.jobs])]).astype(np.int64)
        n = int(self.offsets[-1])
        self.n = n
        dev = torch.device(f"cuda:{backend.device}")
        self.bufs = {
            "Y": torch.empty((n, 8), dtype=torch.float64, device=dev),
            "lam": torch.empty(n, dtype=torch.float64, device=dev),
            "n_steps": torch.empty(n, dtype=torch.int64, device=dev),
            "n_rejected": torch.empty(n, dtype=torch.int64, device=dev),
            "state": torch.empty(n, dtype=torch.int64, device=dev),
            "m0": torch.empty(n, dtype=torch.float64, device=dev),
            "m1": torch.empty(n, dtype=torch.float64, device=dev),
            "m2": torch.empty(n, dtype=torch.float64, device=dev),
            "m3": torch.empty(n, dtype=torch.float64, device=dev),
        }
        o = _lib.KrOutputs()
        o.Y, o.lam = self.bufs["Y"].data_ptr(), self.bufs["lam"].data_ptr()
        o.n_steps, o.n_rejected = self.bufs["n_steps"].data_ptr(), self.bufs["n_rejected"].data_ptr()
        o.state = self.bufs["state"].data_ptr()
        o.max_null_error, o.max_energy_drift = self.bufs["m0"].data_ptr(), self.bufs["m1"].data_ptr()
        o.max_lz_drift, o.max_carter_drift = self.bufs["m2"].data_ptr(), self.bufs["m3"].data_ptr()
        self.outputs = o
        arr = (_lib.KrJob * len(self.jobs))(*[_kr_job(j) for j in self.jobs])
        self.tune = tune or _lib.KrTune(256, 2, 4, 1)
        h = ctypes.c_void_p()
        self.precision = precision
        if precision is None:
            _lib.check(_lib.lib().kr_batch_create(backend.handle.ptr, arr, len(self.jobs), ctypes.byref(self.tune),
                                                  ctypes.byref(o), ctypes.byref(h)), "kr_batch_create")
        else:
            prec, self._keep = _precision_struct(self.jobs, precision, fp32_min_rtol)
            _lib.check(_lib.lib().kr_batch_create_precision(backend.handle.ptr, arr, len(self.jobs),
                                                            ctypes.byref(self.tune), ctypes.byref(prec),
                                                            ctypes.byref(o), ctypes.byref(h)),
                       "kr_batch_create_precision")
        self._b = h

    def run(self, graph: bool = True, stream=None) -> None:
        """Enqueue one run on ``stream`` (default: torch's current stream)."""
        s = stream if stream is not None else self.torch.cuda.current_stream(self.backend.device)
        fn = _lib.lib().kr_batch_graph_launch if graph else _lib.lib().kr_batch_run
        _lib.check(fn(self._b, ctypes.c_void_p(s.cuda_stream)), "kr_batch run")

    def timed(self, repeats: int = 3, graph: bool = True) -> float:
        """Best device time (ms) of ``repeats`` runs, HIP events on the launch stream."""
        torch = self.torch
        s = torch.cuda.current_stream(self.backend.device)
        best = math.inf
        for _ in range(repeats):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            self.run(graph=graph, stream=s)
            e1.record(s)
            e1.synchronize()
            best = min(best, e0.elapsed_time(e1))
        return best

    def check_flags(self) -> None:
        flags = ctypes.c_int32(0)
        _lib.check(_lib.lib().kr_batch_flags(self._b, ctypes.byref(flags)), "kr_batch_flags")
        _lib.raise_for_flags(flags.value)

    def stats(self) -> _lib.KrPlannedStats:
        """Counters of the last run of a precision batch (band photons, repairs, FP64 attempts, flags)."""
        st = _lib.KrPlannedStats()
        _lib.check(_lib.lib().kr_batch_stats(self._b, ctypes.byref(st)), "kr_batch_stats")
        _lib.raise_for_flags(st.flags)
        return st

    def results(self) -> list[BatchResult]:
        """Per-image results (host copies), in job order."""
        self.torch.cuda.synchronize(self.backend.device)
        self.check_flags()
        h = {k: v.cpu().numpy() for k, v in self.bufs.items()}
        out = []
        for k in range(len(self.jobs)):
            a, b = int(self.offsets[k]), int(self.offsets[k + 1])
            out.append(BatchResult(
                Y=h["Y"][a:b], state=h["state"][a:b], n_steps=h["n_steps"][a:b], lam=h["lam"][a:b],
                max_null_error=h["m0"][a:b], max_energy_drift=h["m1"][a:b], max_lz_drift=h["m2"][a:b],
                max_carter_drift=h["m3"][a:b], runtime_s=0.0, n_rejected=h["n_rejected"][a:b],
            ))
        return out

    def close(self) -> None:
        if getattr(self, "_b", None) is not None and self._b.value:
            _lib.lib().kr_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass


def _precision_struct(jobs: list[ImageJob], options, fp32_min_rtol: float) -> tuple[_lib.KrBatchPrecision, tuple]:
    """Each image's analytic edge (Bardeen's curve about its centroid; shrunk by the finite-distance factor
    for a ZAMO camera) for ``kr_batch_create_precision``."""
    from kerrray_amd.boundary import finite_distance_scale
    from kerrray_amd.planned import edge_table

    tables = []
    for job in jobs:
        st = Spacetime(job.mass, job.spin)
        t = edge_table(st, job.inclination_deg)
        k = finite_distance_scale(st, job.r_o) if job.camera == "zamo" else 1.0
        tables.append((t.centre[0] / k, t.centre[1] / k, t.radii / k))
    n_angles = tables[0][2].size
    ca = np.ascontiguousarray([t[0] for t in tables], dtype=np.float64)
    cb = np.ascontiguousarray([t[1] for t in tables], dtype=np.float64)
    radii = np.ascontiguousarray(np.concatenate([t[2] for t in tables]), dtype=np.float64)
    p = _lib.KrBatchPrecision()
    p.band_px, p.n_angles, p.axis_lz = int(options.band_px), int(n_angles), float(options.axis_lz)
    p.fp32_min_rtol = float(fp32_min_rtol)
    p.centre_alpha, p.centre_beta, p.radii = ca.ctypes.data, cb.ctypes.data, radii.ctypes.data
    return p, (ca, cb, radii)


def _cache_path() -> str:
    return os.environ.get("KERRRAY_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "kerrray_amd", "tune.json"))


def autotune(backend: HipBackend, jobs: list[ImageJob], *, block_sizes=BLOCK_SIZES, waves=WAVES_PER_SIMD,
             intervals=REFILL_INTERVALS, repeats: int = 2, cache: bool = True,
             log=None) -> tuple[_lib.KrTune, list[dict]]:
    """Time every (block size, waves per SIMD, refill interval) on ``jobs``, images in cost order (most
    expensive photons first); return the fastest and the table."""
    import torch

    key = None
    if cache:
        import hashlib

        sig = json.dumps([asdict(j) for j in jobs], sort_keys=True)
        digest = hashlib.sha256(f"{sig}|{block_sizes}|{waves}|{intervals}|order1".encode()).hexdigest()[:16]
        key = f"{torch.cuda.get_device_name(backend.device)}|{digest}"
        try:
            with open(_cache_path()) as f:
                hit = json.load(f).get(key)
            if hit:
                return _lib.KrTune(hit["block_size"], hit["waves_per_simd"], hit["refill_interval"], 1), hit["table"]
        except (OSError, ValueError):
            pass
    table = []
    best = None
    for bs, w, k in itertools.product(block_sizes, waves, intervals):
        tune = _lib.KrTune(bs, w, k, 1)
        batch = DeviceBatch(backend, jobs, tune)
        try:
            batch.run(graph=False)
            ms = batch.timed(repeats=repeats, graph=False)
        finally:
            batch.close()
        row = {"block_size": bs, "waves_per_simd": w, "refill_interval": k, "ms": ms}
        table.append(row)
        if log:
            log(f"autotune block={bs} waves={w} refill={k}: {ms:.3f} ms")
        if best is None or ms < best["ms"]:
            best = row
    tune = _lib.KrTune(best["block_size"], best["waves_per_simd"], best["refill_interval"], 1)
    if cache and key:
        try:
            path = _cache_path()
            os.makedirs(os.path.dirname(path), exist_ok=True)
            data = {}
            if os.path.exists(path):
                with open(path) as f:
                    data = json.load(f)
            data[key] = {**best, "table": table, "time": time.time()}
            with open(path, "w") as f:
                json.dump(data, f)
        except (OSError, ValueError):
            pass
    return tune, table
