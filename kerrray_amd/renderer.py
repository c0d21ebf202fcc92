"""Accretion-disk images: ``render_disk`` on the device.

Mirrors ``DiskParams``, ``DiskModel``, ``event_options``, ``DiskImage``,
``render_disk``, ``save_disk_image`` and ``load_disk_image_arrays`` of the
reference (src/kerrray/raytracing/renderer.py:108-260,
src/kerrray/physics/accretion.py:77-205). One call (``kr_render_disk``)
builds every pixel's photon on the device (ZAMO camera, or Bardeen image
plane), integrates it with the disk-plane event (``DISK_HIT`` where the ray
crosses theta = pi/2 inside [r_in, r_out], numba_kernel.py:439-454) and shades
each crossing in a second kernel: the redshift factor of the Keplerian gas,

    g = (-p_t) / (-p_mu u^mu),  u = u^t (1, 0, 0, Omega),  Omega = +-sqrt(M) / (r^1.5 +- |a| sqrt(M))

(physics/redshift.py:63-151, BPT 1972 eq. 2.16) and ``I_obs = g^n (r / r_in)^(-p)``
on the annulus (accretion.py:180-196). Pixels that miss the disk get
``I_obs = 0`` and ``g = r_hit = NaN``.
"""

from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass
from pathlib import Path
from typing import Any

import numpy as np
from numpy.typing import ArrayLike, NDArray

from kerrray_amd import _lib
from kerrray_amd.backends import Backend, HipBackend, HostBatch, get_backend, kernel_params
from kerrray_amd.camera import BardeenCamera, Camera, image_struct
from kerrray_amd.geodesics import EventOptions, IntegratorOptions, TerminationOptions
from kerrray_amd.geometry import Spacetime
from kerrray_amd.orbits import equatorial_photon_orbit_radius, isco_radius
from kerrray_amd.photons import BatchResult, TerminationState
from kerrray_amd.rays import ProgressCallback

__all__ = [
    "INTENSITY_LAWS", "DiskImage", "DiskModel", "DiskParams", "event_options", "load_disk_image_arrays",
    "redshift_factor", "render_disk", "save_disk_image",
]

INTENSITY_LAWS = {"g3": 3, "g4": 4}
"""Power of ``g`` multiplying the emitted intensity for each ``intensity_law``."""

FloatArray = NDArray[np.float64]


@dataclass(frozen=True)
class DiskParams:
    """The ``disk`` parameter block (accretion.py:77-110); ``r_in = None`` selects the ISCO."""

    enabled: bool = False
    r_in: float | None = None
    r_out: float = 20.0
    emissivity_index: float = 3.0
    intensity_law: str = "g4"
    prograde: bool = True

    def __post_init__(self) -> None:
        if not math.isfinite(self.r_out) or self.r_out <= 0.0:
            raise ValueError(f"'disk.r_out' must be finite and > 0, got {self.r_out!r}")
        if self.r_in is not None and not (math.isfinite(self.r_in) and 0.0 < self.r_in < self.r_out):
            raise ValueError(f"'disk.r_in' must satisfy 0 < r_in < r_out, got {self.r_in!r}")
        if not math.isfinite(self.emissivity_index):
            raise ValueError(f"'disk.emissivity_index' must be finite, got {self.emissivity_index!r}")
        if self.intensity_law not in INTENSITY_LAWS:
            raise ValueError(f"'disk.intensity_law' must be one of {', '.join(sorted(INTENSITY_LAWS))}, "
                             f"got {self.intensity_law!r}")


@dataclass(frozen=True)
class DiskModel:
    """Resolved disk (accretion.py:117-177): annulus, emissivity index, intensity law, sense of rotation."""

    r_in: float
    r_out: float
    emissivity_index: float = 3.0
    intensity_law: str = "g4"
    prograde: bool = True

    def __post_init__(self) -> None:
        for name in ("r_in", "r_out", "emissivity_index"):
            if not math.isfinite(float(getattr(self, name))):
                raise ValueError(f"{name} must be finite, got {getattr(self, name)!r}")
        if not (0.0 < self.r_in < self.r_out):
            raise ValueError(f"need 0 < r_in < r_out, got r_in={self.r_in!r}, r_out={self.r_out!r}")
        if self.intensity_law not in INTENSITY_LAWS:
            raise ValueError(f"intensity_law must be one of {sorted(INTENSITY_LAWS)}, got {self.intensity_law!r}")

    @property
    def redshift_power(self) -> int:
        return INTENSITY_LAWS[self.intensity_law]

    @classmethod
    def from_params(cls, st: Spacetime, params: DiskParams) -> DiskModel:
        """``r_in = None`` becomes the ISCO; the inner edge must lie outside the circular photon orbit."""
        r_in = params.r_in if params.r_in is not None else isco_radius(st, prograde=params.prograde)
        r_ph = equatorial_photon_orbit_radius(st, prograde=params.prograde)
        if not r_in > r_ph:
            raise ValueError(f"disk r_in = {r_in:.6g} must exceed the circular photon orbit radius r_ph = {r_ph:.6g} "
                             "(no circular gas orbit below it)")
        return cls(r_in=float(r_in), r_out=float(params.r_out), emissivity_index=float(params.emissivity_index),
                   intensity_law=params.intensity_law, prograde=bool(params.prograde))

    def kr(self) -> _lib.KrDisk:
        return _lib.KrDisk(float(self.r_in), float(self.r_out), float(self.emissivity_index), self.redshift_power,
                           1 if self.prograde else 0)


def event_options(disk: DiskModel) -> EventOptions:
    """The disk-plane event of this disk (``DISK_HIT`` on ``r_in <= r <= r_out``)."""
    return EventOptions(disk_plane=True, r_in=disk.r_in, r_out=disk.r_out)


@dataclass
class DiskImage:
    """Rendered disk image and the per-pixel quantities behind it (renderer.py:108-160), shape ``(H, W)``."""

    intensity: FloatArray
    g: FloatArray
    hit: NDArray[np.bool_]
    r_hit: FloatArray
    state: NDArray[np.int64]
    alpha: FloatArray
    beta: FloatArray
    hit_states: FloatArray
    result: BatchResult
    camera: Camera | BardeenCamera
    spacetime: Spacetime
    disk: DiskModel
    backend: str
    kernel_ms: float = 0.0

    @property
    def shape(self) -> tuple[int, int]:
        return self.camera.shape

    @property
    def n_hits(self) -> int:
        return int(np.count_nonzero(self.hit))

    def counts(self) -> dict[str, int]:
        """Pixels per termination-state name (states with no pixel omitted)."""
        out: dict[str, int] = {}
        for member in TerminationState:
            n = int(np.count_nonzero(self.state == int(member)))
            if n:
                out[member.name] = n
        return out


def render_disk(
    st: Spacetime,
    cam: Camera | BardeenCamera,
    integ: IntegratorOptions,
    term: TerminationOptions,
    disk: DiskModel,
    *,
    backend: str | Backend = "hip",
    progress: ProgressCallback | None = None,
) -> DiskImage:
    """Render the disk seen by ``cam`` (renderer.py:163-216): one device trace with the disk event, then shading."""
    engine = get_backend(backend) if isinstance(backend, str) else backend
    if not isinstance(engine, HipBackend):
        raise TypeError("render_disk runs on the hip backend")
    if integ.method == "dop853":
        raise ValueError("method 'dop853' is available in integrate() only; the hip backend supports 'rk4' and 'rk45'")
    t0 = time.perf_counter()
    params = kernel_params(st, integ, term, event_options(disk))
    img = image_struct(st, cam)
    n = cam.n_rays
    out = HostBatch(n, integ.np_dtype, True)
    intensity, g, r_hit = np.empty(n), np.empty(n), np.empty(n)
    kd = disk.kr()
    ms = ctypes.c_double(0.0)
    with engine.handle._lock:
        _lib.check(_lib.lib().kr_render_disk(engine.handle.ptr, ctypes.byref(params), ctypes.byref(img),
                                             ctypes.byref(kd), ctypes.byref(out.outputs), intensity.ctypes.data,
                                             g.ctypes.data, r_hit.ctypes.data, ctypes.byref(ms)),
                   "kr_render_disk")
    result = out.result(time.perf_counter() - t0)
    if progress is not None:
        progress(1.0)
    hit = np.asarray(result.event_hit, dtype=bool)
    states = np.full((n, 8), np.nan)
    states[hit] = np.asarray(result.event_Y, dtype=np.float64)[hit]
    shape = cam.shape
    alpha, beta = cam.pixel_coordinates()
    return DiskImage(
        intensity=intensity.reshape(shape), g=g.reshape(shape), hit=hit.reshape(shape), r_hit=r_hit.reshape(shape),
        state=np.asarray(result.state, dtype=np.int64).reshape(shape), alpha=alpha, beta=beta,
        hit_states=states.reshape(shape + (8,)), result=result, camera=cam, spacetime=st, disk=disk,
        backend=engine.name, kernel_ms=ms.value,
    )


def redshift_factor(st: Spacetime, y_hit: ArrayLike, *, prograde: bool = True, backend: Any = "hip") -> FloatArray:
    """``g = nu_obs / nu_em`` of equatorial crossing states for Keplerian gas (redshift.py:131-151), on the device.

    Raises ``ValueError`` where the reference does: a non-finite or
    non-equatorial state, no circular orbit at its radius, or ``g <= 0``.
    """
    import torch

    y = np.asarray(y_hit, dtype=np.float64)
    if y.shape[-1:] != (8,):
        raise ValueError(f"state must end in a dimension of size 8, got {y.shape}")
    engine = get_backend(backend) if isinstance(backend, str) else backend
    dev = torch.device("cuda", engine.handle.device)
    rows = torch.as_tensor(np.ascontiguousarray(y.reshape(-1, 8)), device=dev)
    n = rows.shape[0]
    hit = torch.ones(n, dtype=torch.uint8, device=dev)
    out = torch.empty((3, n), dtype=torch.float64, device=dev)
    flags = torch.zeros(1, dtype=torch.int32, device=dev)
    kd = _lib.KrDisk(1.0, 2.0, 0.0, 4, 1 if prograde else 0)
    stream = torch.cuda.current_stream(dev).cuda_stream
    with engine.handle._lock:
        _lib.check(_lib.lib().kr_disk_shade_dev(engine.handle.ptr, ctypes.byref(kd), st.mass, st.spin,
                                                _lib.KR_FLOAT64, rows.data_ptr(), hit.data_ptr(), n,
                                                out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr(),
                                                flags.data_ptr(), ctypes.c_void_p(stream)), "kr_disk_shade_dev")
    g = out[1].cpu().numpy()
    _lib.raise_for_flags(int(flags.item()))
    return g.reshape(y.shape[:-1])


def save_disk_image(path: Path | str, image: DiskImage) -> Path:
    """Compressed ``.npz`` with the reference's keys (renderer.py:219-255)."""
    target = Path(path)
    target.parent.mkdir(parents=True, exist_ok=True)
    cam, disk, st = image.camera, image.disk, image.spacetime
    np.savez_compressed(
        target, intensity=image.intensity, g=image.g, hit=image.hit, r_hit=image.r_hit, state=image.state,
        alpha=image.alpha, beta=image.beta, hit_states=image.hit_states,
        lam=np.asarray(image.result.lam, dtype=np.float64), n_steps=np.asarray(image.result.n_steps),
        max_null_error=np.asarray(image.result.max_null_error, dtype=np.float64), mass=st.mass, spin=st.spin,
        inclination_deg=cam.inclination_deg, camera_radius=cam.radius, fov=cam.fov, resolution=cam.resolution,
        r_in=disk.r_in, r_out=disk.r_out, emissivity_index=disk.emissivity_index,
        redshift_power=disk.redshift_power, prograde=disk.prograde,
    )
    return target


def load_disk_image_arrays(path: Path | str) -> dict[str, Any]:
    """Read a file written by :func:`save_disk_image` (pickle-free) into a ``dict`` of arrays and scalars."""
    with np.load(Path(path), allow_pickle=False) as data:
        return {key: (data[key].item() if data[key].ndim == 0 else data[key]) for key in data.files}
