"""Shadow images: trace every pixel, classify, keep the captured mask.

Mirrors ``ShadowImage``, ``state_counts`` and ``compute_shadow`` of the
reference (src/kerrray/raytracing/shadow.py:59-211), including its ``.npz``
file layout (format version 1). The camera is the reference's ZAMO
:class:`~kerrray_amd.camera.Camera` or a
:class:`~kerrray_amd.camera.BardeenCamera`. On the ``hip`` backend the initial
states of either are generated on the device (``kr_trace_image``: one pixel
per lane, 8 x 8 pixel tiles per wavefront) instead of being built on the host
and copied over.
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass
from pathlib import Path
from typing import Any

import numpy as np
from numpy.typing import NDArray

from kerrray_amd import _lib
from kerrray_amd.backends import Backend, HipBackend, HostBatch, get_backend, kernel_params
from kerrray_amd.camera import BardeenCamera, Camera, bardeen_initial_states, image_struct, initial_states
from kerrray_amd.geodesics import IntegratorOptions, TerminationOptions
from kerrray_amd.geometry import Spacetime
from kerrray_amd.photons import FAILURE_STATES, BatchResult, TerminationState
from kerrray_amd.rays import DEFAULT_CHUNK_SIZE, ProgressCallback, trace_rays

__all__ = ["NPZ_FORMAT_VERSION", "ShadowImage", "compute_shadow", "state_counts", "trace_image"]

NPZ_FORMAT_VERSION = 1
_CAMERA_FIELDS = ("radius", "inclination_deg", "phi_deg", "fov", "resolution")
_RESULT_ARRAYS = (
    "Y", "state", "n_steps", "lam", "max_null_error", "max_energy_drift", "max_lz_drift", "max_carter_drift",
)
_OPTIONAL_RESULT_ARRAYS = ("event_Y", "event_hit", "n_rejected")


def state_counts(state: NDArray[np.integer]) -> dict[str, int]:
    """Pixels per termination-state name (every member, zeros included)."""
    codes = np.asarray(state)
    return {member.name: int(np.count_nonzero(codes == int(member))) for member in TerminationState}


@dataclass
class ShadowImage:
    alpha: NDArray[np.float64]
    beta: NDArray[np.float64]
    state: NDArray[np.int64]
    captured: NDArray[np.bool_]
    counts: dict[str, int]
    result: BatchResult
    camera: Camera | BardeenCamera
    spacetime: Spacetime

    @property
    def shape(self) -> tuple[int, int]:
        return self.camera.shape

    @property
    def n_rays(self) -> int:
        return self.camera.n_rays

    @property
    def pixel_size(self) -> float:
        return self.camera.pixel_size

    def count(self, state: TerminationState) -> int:
        return self.counts[state.name]

    @property
    def n_failed(self) -> int:
        return sum(self.count(s) for s in FAILURE_STATES)

    @property
    def n_other(self) -> int:
        settled = self.count(TerminationState.CAPTURED) + self.count(TerminationState.ESCAPED)
        return self.n_rays - settled - self.n_failed

    def save(self, path: Path | str) -> Path:
        """Compressed ``.npz`` in the reference's version-1 layout (shadow.py:119-142)."""
        target = Path(path)
        target.parent.mkdir(parents=True, exist_ok=True)
        payload: dict[str, Any] = {
            "format_version": np.int64(NPZ_FORMAT_VERSION),
            "alpha": self.alpha,
            "beta": self.beta,
            "image_state": self.state,
            "captured": self.captured,
            "mass": np.float64(self.spacetime.mass),
            "spin": np.float64(self.spacetime.spin),
            "runtime_s": np.float64(self.result.runtime_s),
        }
        for name in _CAMERA_FIELDS:
            payload[f"camera_{name}"] = np.asarray(getattr(self.camera, name))
        if isinstance(self.camera, BardeenCamera):
            payload["camera_model"] = np.asarray("bardeen")
        for name in _RESULT_ARRAYS:
            payload[f"result_{name}"] = np.asarray(getattr(self.result, name))
        for name in _OPTIONAL_RESULT_ARRAYS:
            value = getattr(self.result, name)
            if value is not None:
                payload[f"result_{name}"] = np.asarray(value)
        np.savez_compressed(target, **payload)
        return target

    @classmethod
    def load(cls, path: Path | str) -> ShadowImage:
        """Read a file written by :meth:`save` (pickle-free)."""
        with np.load(Path(path), allow_pickle=False) as data:
            version = int(data["format_version"])
            if version != NPZ_FORMAT_VERSION:
                raise ValueError(f"unsupported shadow file version {version} (expected {NPZ_FORMAT_VERSION})")
            # Files without a camera model are the reference's, whose camera is the ZAMO one.
            model = str(data["camera_model"]) if "camera_model" in data.files else "zamo"
            camera = (BardeenCamera if model == "bardeen" else Camera)(
                radius=float(data["camera_radius"]),
                inclination_deg=float(data["camera_inclination_deg"]),
                phi_deg=float(data["camera_phi_deg"]),
                fov=float(data["camera_fov"]),
                resolution=int(data["camera_resolution"]),
            )
            spacetime = Spacetime(mass=float(data["mass"]), spin=float(data["spin"]))
            fields: dict[str, Any] = {name: data[f"result_{name}"] for name in _RESULT_ARRAYS}
            for name in _OPTIONAL_RESULT_ARRAYS:
                key = f"result_{name}"
                fields[name] = data[key] if key in data.files else None
            result = BatchResult(runtime_s=float(data["runtime_s"]), **fields)
            state = np.asarray(data["image_state"], dtype=np.int64)
            return cls(
                alpha=np.asarray(data["alpha"], dtype=np.float64),
                beta=np.asarray(data["beta"], dtype=np.float64),
                state=state,
                captured=np.asarray(data["captured"], dtype=bool),
                counts=state_counts(state),
                result=result,
                camera=camera,
                spacetime=spacetime,
            )


def kr_image(st: Spacetime, r_o: float, inclination_deg: float, fov: float = 1.0, resolution: int = 2) -> _lib.KrImage:
    """The ``kr_image`` of a Bardeen image plane (image points only: fov and resolution are placeholders)."""
    import math

    img = _lib.KrImage()
    img.mass = st.mass
    img.spin = st.spin
    img.r_o = float(r_o)
    img.theta_o = math.radians(inclination_deg)
    img.fov = float(fov)
    img.resolution = int(resolution)
    return img


def trace_image(engine: HipBackend, st: Spacetime, cam: Camera | BardeenCamera, integ: IntegratorOptions,
                term: TerminationOptions) -> BatchResult:
    """All pixels of ``cam`` on the device in one launch (``kr_trace_image``), row-major."""
    if integ.method == "dop853":
        raise ValueError("method 'dop853' is available in integrate() only; the hip backend supports 'rk4' and 'rk45'")
    t0 = time.perf_counter()
    params = kernel_params(st, integ, term, None)
    img = image_struct(st, cam)
    out = HostBatch(cam.n_rays, integ.np_dtype, False)
    ms = ctypes.c_double(0.0)
    with engine.handle._lock:
        _lib.check(
            _lib.lib().kr_trace_image(engine.handle.ptr, ctypes.byref(params), ctypes.byref(img),
                                      ctypes.byref(out.outputs), ctypes.byref(ms)),
            "kr_trace_image",
        )
    engine.last_kernel_ms = ms.value
    return out.result(time.perf_counter() - t0)


def compute_shadow(
    st: Spacetime,
    cam: Camera | BardeenCamera,
    integ: IntegratorOptions,
    term: TerminationOptions,
    *,
    backend: str | Backend = "hip",
    progress: ProgressCallback | None = None,
    chunk_size: int = DEFAULT_CHUNK_SIZE,
    schedule: str = "auto",
) -> ShadowImage:
    """Trace every pixel of ``cam`` and classify it (shadow.py:179-211).

    On the ``hip`` backend the device generates and traces every pixel's
    photon. With ``schedule="auto"`` an FP64 Dormand-Prince image runs through
    the planned persistent kernel (the edge band's long photons first; every
    row bit for bit the plain kernel's, about 1.4x faster at 512^2), and
    anything else (RK4, float32) through the plain kernel, one thread per
    pixel; ``schedule="plain"`` always takes the plain kernel. Another
    :class:`Backend` gets the initial states as an (N, 8) array and traces
    them in chunks through :func:`trace_rays`.
    """
    if not isinstance(cam, (Camera, BardeenCamera)):
        raise TypeError("compute_shadow needs a Camera or a BardeenCamera")
    if schedule not in ("auto", "plain"):
        raise ValueError(f"unknown schedule {schedule!r}: use 'auto' or 'plain'")
    engine = get_backend(backend) if isinstance(backend, str) else backend
    if isinstance(engine, HipBackend):
        if schedule == "auto" and integ.method == "rk45" and integ.np_dtype == np.float64:
            from kerrray_amd.planned import PlannedOptions, trace_planned

            result = trace_planned(engine, st, cam, integ, term, PlannedOptions(fp32_bulk=False))[0]
        else:
            result = trace_image(engine, st, cam, integ, term)
        if progress is not None:
            progress(1.0)
    else:
        Y0 = initial_states(cam, st) if isinstance(cam, Camera) else bardeen_initial_states(cam, st)
        result = trace_rays(st, Y0, integ, term, backend=engine, progress=progress, chunk_size=chunk_size)
    alpha, beta = cam.pixel_coordinates()
    state = np.asarray(result.state, dtype=np.int64).reshape(cam.shape)
    return ShadowImage(
        alpha=alpha, beta=beta, state=state, captured=state == int(TerminationState.CAPTURED),
        counts=state_counts(state), result=result, camera=cam, spacetime=st,
    )
