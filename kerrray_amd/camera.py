"""Cameras: the ZAMO observer and the image plane with Bardeen initial conditions.

:class:`Camera` is the reference's observer camera
(src/kerrray/raytracing/camera.py:100-381): a zero-angular-momentum observer
(ZAMO) at ``(r_o, theta_o, phi_o)`` whose pixel ``(alpha, beta)`` receives
the photon with tetrad components ``p^(a) = (1, sqrt(1 - rho^2), beta / r_o,
-alpha / r_o)``, ``rho^2 = (alpha^2 + beta^2) / r_o^2``, projected onto
Boyer-Lindquist momenta, reversed for backward tracing and re-normalised to
be exactly null (``_null_pt``, camera.py:265-295). The device kernels build
these states per pixel (``kr_image_states_dev``, and every image entry point
for an image whose camera is ``KR_CAMERA_ZAMO``); :func:`initial_states`
returns them as the reference's function does.

:class:`BardeenCamera` has the fields and pixel grid of the reference's
``Camera`` (src/kerrray/raytracing/camera.py:100-222): ``resolution`` x
``resolution`` pixel centres ``alpha_j = fov (2 j + 1 - n) / n``,
``beta_i = -alpha_i``, row 0 on top. Each pixel's photon is built from the
celestial constants of its image point, as the reference's
``_off_equatorial_state`` does (src/kerrray/benchmarks/rayset.py:158-165,
src/kerrray/geodesics/initial_conditions.py:50-112)::

    xi  = -alpha sin(theta_o)                       (= L_z / E)
    eta = beta^2 - a^2 cos^2(theta_o) + xi^2 cot^2(theta_o)   (= Q / E^2)
    y0  = [0, r_o, theta_o, 0, -E, -sqrt(R(r_o))/Delta, s_theta sqrt(Theta(theta_o)), xi],  E = 1

(forward in lambda, moving inward). ``s_theta = -1`` for ``beta >= 0`` and
``+1`` otherwise, the direction in which the backward ray of a pixel above
the image centre leaves the observer. Capture depends only on
``(xi, eta)``, so the mask does not depend on ``s_theta``.

The device kernels generate these states themselves (one pixel per lane);
:func:`bardeen_initial_states` builds the same states on the host, e.g. to
feed :meth:`Backend.integrate_batch`.
"""

from __future__ import annotations

import ctypes
import logging
import math
from dataclasses import dataclass
from typing import Any

import numpy as np
from numpy.typing import ArrayLike, NDArray

from kerrray_amd.geometry import Spacetime, outer_horizon

__all__ = [
    "INCLINATION_EPSILON_DEG", "NULL_CHECK_TOLERANCE", "RADICAND_TOLERANCE", "BardeenCamera", "Camera",
    "bardeen_initial_states", "bardeen_points", "image_plane_states", "image_plane_states_device", "image_struct",
    "initial_states", "zamo_tetrad",
]

INCLINATION_EPSILON_DEG = 1e-3
RADICAND_TOLERANCE = 1e-12
NULL_CHECK_TOLERANCE = 1e-10
"""Largest ``abs(g^{mu nu} p_mu p_nu)`` of a ZAMO photon before re-normalisation (camera.py:85-96)."""

logger = logging.getLogger("kerrray_amd")


def _checked_fields(radius: Any, inclination_deg: Any, phi_deg: Any, fov: Any,
                    resolution: Any) -> tuple[float, float, float, float, int]:
    """The validation and axis clamp of ``Camera.__post_init__`` (camera.py:131-168)."""
    radius_f, inclination, phi, fov_f = float(radius), float(inclination_deg), float(phi_deg), float(fov)
    if not (math.isfinite(radius_f) and radius_f > 0.0):
        raise ValueError(f"radius must be a finite number > 0, got {radius!r}")
    if not (math.isfinite(inclination) and 0.0 <= inclination <= 180.0):
        raise ValueError(f"inclination_deg must lie in [0, 180], got {inclination_deg!r}")
    if not math.isfinite(phi):
        raise ValueError(f"phi_deg must be finite, got {phi_deg!r}")
    if not (math.isfinite(fov_f) and fov_f > 0.0):
        raise ValueError(f"fov must be a finite number > 0, got {fov!r}")
    if isinstance(resolution, bool) or not isinstance(resolution, (int, np.integer)):
        raise ValueError(f"resolution must be an integer, got {resolution!r}")
    res = int(resolution)
    if res < 2 or res % 2 != 0:
        raise ValueError(f"resolution must be an even integer >= 2 so that alpha = 0 is never sampled, got {res!r}")
    low, high = INCLINATION_EPSILON_DEG, 180.0 - INCLINATION_EPSILON_DEG
    clamped = min(max(inclination, low), high)
    if clamped != inclination:
        logger.warning("camera inclination_deg = %g lies on the spin axis, where the ZAMO tetrad is singular; "
                       "clamped to %g deg", inclination, clamped)
    return radius_f, clamped, phi, fov_f, res


@dataclass(frozen=True)
class BardeenCamera:
    """Observer at ``(radius, inclination_deg)`` with a square image plane of half-width ``fov``."""

    radius: float
    inclination_deg: float
    phi_deg: float = 0.0
    fov: float = 12.0
    resolution: int = 64

    def __post_init__(self) -> None:
        radius = float(self.radius)
        inclination = float(self.inclination_deg)
        fov = float(self.fov)
        if not (math.isfinite(radius) and radius > 0.0):
            raise ValueError(f"radius must be a finite number > 0, got {self.radius!r}")
        if not (math.isfinite(inclination) and 0.0 <= inclination <= 180.0):
            raise ValueError(f"inclination_deg must lie in [0, 180], got {self.inclination_deg!r}")
        if float(self.phi_deg) != 0.0:
            raise ValueError("the Bardeen camera sits at phi_o = 0 (rayset.py:158-165); got phi_deg != 0")
        if not (math.isfinite(fov) and fov > 0.0):
            raise ValueError(f"fov must be a finite number > 0, got {self.fov!r}")
        res = self.resolution
        if isinstance(res, bool) or not isinstance(res, (int, np.integer)):
            raise ValueError(f"resolution must be an integer, got {res!r}")
        res = int(res)
        if res < 2 or res % 2 != 0:
            raise ValueError(f"resolution must be an even integer >= 2, got {res!r}")
        low, high = INCLINATION_EPSILON_DEG, 180.0 - INCLINATION_EPSILON_DEG
        object.__setattr__(self, "radius", radius)
        object.__setattr__(self, "inclination_deg", min(max(inclination, low), high))
        object.__setattr__(self, "phi_deg", 0.0)
        object.__setattr__(self, "fov", fov)
        object.__setattr__(self, "resolution", res)

    @property
    def n_rays(self) -> int:
        return self.resolution * self.resolution

    @property
    def shape(self) -> tuple[int, int]:
        return (self.resolution, self.resolution)

    @property
    def pixel_size(self) -> float:
        return 2.0 * self.fov / self.resolution

    @property
    def inclination_rad(self) -> float:
        return math.radians(self.inclination_deg)

    @property
    def phi_rad(self) -> float:
        return 0.0

    def pixel_coordinates(self) -> tuple[NDArray[np.float64], NDArray[np.float64]]:
        """``(alpha, beta)`` of the pixel centres, shape ``(H, W)`` (camera.py:206-222)."""
        n = self.resolution
        centres = self.fov * (2.0 * np.arange(n, dtype=np.float64) + 1.0 - n) / n
        alpha, beta = np.meshgrid(centres, -centres)
        return alpha, beta


def bardeen_points(
    st: Spacetime, r_o: float, inclination_deg: float, alpha: ArrayLike, beta: ArrayLike
) -> NDArray[np.float64]:
    """Bardeen photons (N, 8) for image points ``(alpha, beta)`` (module docstring).

    Vectorised over the points with the reference's scalar expressions;
    NumPy's elementwise ``+ - * / sqrt`` round like Python floats, and the
    per-image factors use Python's own ``math`` and ``**``. Raises
    ``ValueError`` where ``photon_from_constants`` raises.
    """
    if not r_o > outer_horizon(st):
        raise ValueError(f"r = {r_o!r} must lie outside the outer horizon r_plus = {outer_horizon(st)!r}")
    th = math.radians(inclination_deg)
    sin_o, cos_o = math.sin(th), math.cos(th)
    a, m = st.a, st.mass
    alpha = np.asarray(alpha, dtype=np.float64).ravel()
    beta = np.asarray(beta, dtype=np.float64).ravel()
    a2c2 = a**2 * cos_o**2
    cot2 = (cos_o / sin_o) ** 2
    xi = -alpha * sin_o
    eta = beta * beta - a2c2 + xi * xi * cot2
    dlt = r_o * r_o - 2.0 * m * r_o + a**2
    E = 1.0
    term1 = E * (r_o * r_o + a * a) - a * xi
    d = xi - a * E
    term2 = dlt * (eta + d * d)
    big_r = term1 * term1 - term2
    scale_r = term1 * term1 + np.abs(term2)
    c2, s2 = math.cos(th) ** 2, math.sin(th) ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        lz_term = np.where(xi == 0.0, 0.0, xi * xi / s2)
    aaee = a * a * E * E
    big_theta = eta - c2 * (lz_term - aaee)
    scale_theta = np.abs(eta) + c2 * (lz_term + aaee)
    if np.any(big_r < -RADICAND_TOLERANCE * np.maximum(scale_r, 1.0)):
        raise ValueError("R(r) < 0: the requested constants of motion are inconsistent here")
    if np.any(big_theta < -RADICAND_TOLERANCE * np.maximum(scale_theta, 1.0)):
        raise ValueError("Theta(theta) < 0: the requested constants of motion are inconsistent here")
    y = np.empty((alpha.size, 8))
    y[:, 0] = 0.0
    y[:, 1] = r_o
    y[:, 2] = th
    y[:, 3] = 0.0
    y[:, 4] = -E
    y[:, 5] = -1.0 * np.sqrt(np.maximum(big_r, 0.0)) / dlt
    y[:, 6] = np.where(beta >= 0.0, -1.0, 1.0) * np.sqrt(np.maximum(big_theta, 0.0))
    y[:, 7] = xi
    return y


def bardeen_initial_states(cam: BardeenCamera, st: Spacetime) -> NDArray[np.float64]:
    """Row-major (n*n, 8) initial states of every pixel of ``cam``."""
    alpha, beta = cam.pixel_coordinates()
    return bardeen_points(st, cam.radius, cam.inclination_deg, alpha, beta)


@dataclass(frozen=True)
class Camera:
    """Pinhole camera of a ZAMO at ``(radius, inclination_deg, phi_deg)`` (camera.py:100-222).

    ``inclination_deg`` is clamped into ``[1e-3, 180 - 1e-3]`` degrees with a
    logged warning: the ZAMO tetrad is singular on the spin axis.
    ``resolution`` is even, so ``alpha = 0`` and ``beta = 0`` are never
    sampled.
    """

    radius: float
    inclination_deg: float
    phi_deg: float = 0.0
    fov: float = 12.0
    resolution: int = 64

    def __post_init__(self) -> None:
        fields = _checked_fields(self.radius, self.inclination_deg, self.phi_deg, self.fov, self.resolution)
        for name, value in zip(("radius", "inclination_deg", "phi_deg", "fov", "resolution"), fields):
            object.__setattr__(self, name, value)

    @property
    def n_rays(self) -> int:
        return self.resolution * self.resolution

    @property
    def shape(self) -> tuple[int, int]:
        return (self.resolution, self.resolution)

    @property
    def pixel_size(self) -> float:
        return 2.0 * self.fov / self.resolution

    @property
    def inclination_rad(self) -> float:
        return math.radians(self.inclination_deg)

    @property
    def phi_rad(self) -> float:
        return math.radians(self.phi_deg)

    def pixel_coordinates(self) -> tuple[NDArray[np.float64], NDArray[np.float64]]:
        """``(alpha, beta)`` of the pixel centres, shape ``(H, W)`` (camera.py:206-222)."""
        n = self.resolution
        centres = self.fov * (2.0 * np.arange(n, dtype=np.float64) + 1.0 - n) / n
        alpha, beta = np.meshgrid(centres, -centres)
        return alpha, beta


def image_struct(st: Spacetime, cam: Camera | BardeenCamera) -> Any:
    """The ``kr_image`` of ``cam`` in ``st``: its camera kind, observer and pixel grid."""
    from kerrray_amd import _lib

    img = _lib.KrImage()
    img.mass = st.mass
    img.spin = st.spin
    img.r_o = float(cam.radius)
    img.theta_o = cam.inclination_rad
    img.fov = float(cam.fov)
    img.resolution = int(cam.resolution)
    img.camera = _lib.KR_CAMERA_ZAMO if isinstance(cam, Camera) else _lib.KR_CAMERA_BARDEEN
    img.phi_o = cam.phi_rad
    return img


def initial_states(cam: Camera, st: Spacetime, *, reverse: bool = True, backend: Any = "hip") -> NDArray[np.float64]:
    """Null photon states of every pixel, shape ``(N, 8)``, row-major over ``(H, W)`` (camera.py:298-381).

    Built on the device (``kr_image_states_dev``, one pixel per lane) and
    copied back. ``reverse=True`` (what the tracer integrates) gives
    ``E = -p_t < 0`` and ``p_r < 0``; ``reverse=False`` the physical,
    future-directed photon.

    Raises:
        ValueError: the camera is not outside the outer horizon, or a pixel
            has ``alpha^2 + beta^2 >= r_o^2``.
        RuntimeError: the tetrad projection is not null to
            :data:`NULL_CHECK_TOLERANCE`, or no null ``p_t`` exists.
    """
    import torch

    from kerrray_amd import _lib
    from kerrray_amd.backends import get_backend

    if not isinstance(cam, Camera):
        raise TypeError("initial_states needs a ZAMO Camera; use bardeen_initial_states for a BardeenCamera")
    if cam.radius <= outer_horizon(st):
        raise ValueError(f"camera radius {cam.radius:.6g} must exceed the outer horizon r_plus = {outer_horizon(st):.6g}")
    r_ergo = st.mass + math.sqrt(st.mass * st.mass - st.a**2 * math.cos(cam.inclination_rad) ** 2)
    if cam.radius <= r_ergo:
        logger.warning("camera radius %g lies inside the ergosphere (r_E(theta_o) = %g): E = -p_t may have either "
                       "sign for the physical photon", cam.radius, r_ergo)
    engine = get_backend(backend) if isinstance(backend, str) else backend
    dev = torch.device("cuda", engine.handle.device)
    y = torch.empty((cam.n_rays, 8), dtype=torch.float64, device=dev)
    flags = torch.zeros(1, dtype=torch.int32, device=dev)
    img = image_struct(st, cam)
    stream = torch.cuda.current_stream(dev).cuda_stream
    with engine.handle._lock:
        _lib.check(_lib.lib().kr_image_states_dev(engine.handle.ptr, ctypes.byref(img), 1 if reverse else 0,
                                                  ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(flags.data_ptr()),
                                                  ctypes.c_void_p(stream)), "kr_image_states_dev")
    out = y.cpu().numpy()
    _lib.raise_for_flags(int(flags.item()))
    return out


def image_plane_states_device(cam: Camera, st: Spacetime, alpha, beta, *, reverse: bool = True,
                              backend: Any = "hip"):
    """Null states (N, 8) of arbitrary image points as a device tensor (``kr_point_states_dev``).

    ``alpha``, ``beta`` are float64 device tensors or arrays of N points; the
    construction is that of :func:`initial_states`, so pixel centres give its
    states bit for bit. Flags raise as in :func:`initial_states`.
    """
    import torch

    from kerrray_amd import _lib
    from kerrray_amd.backends import get_backend

    if not isinstance(cam, Camera):
        raise TypeError("image_plane_states needs a ZAMO Camera")
    if cam.radius <= outer_horizon(st):
        raise ValueError(f"camera radius {cam.radius:.6g} must exceed the outer horizon r_plus = {outer_horizon(st):.6g}")
    engine = get_backend(backend) if isinstance(backend, str) else backend
    dev = torch.device("cuda", engine.handle.device)
    a = torch.as_tensor(alpha, dtype=torch.float64, device=dev).reshape(-1).contiguous()
    b = torch.as_tensor(beta, dtype=torch.float64, device=dev).reshape(-1).contiguous()
    if a.shape != b.shape:
        raise ValueError("alpha and beta must have the same shape")
    y = torch.empty((a.numel(), 8), dtype=torch.float64, device=dev)
    flags = torch.zeros(1, dtype=torch.int32, device=dev)
    img = image_struct(st, cam)
    stream = torch.cuda.current_stream(dev).cuda_stream
    with engine.handle._lock:
        _lib.check(_lib.lib().kr_point_states_dev(engine.handle.ptr, ctypes.byref(img), 1 if reverse else 0,
                                                  ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()),
                                                  a.numel(), ctypes.c_void_p(y.data_ptr()),
                                                  ctypes.c_void_p(flags.data_ptr()), ctypes.c_void_p(stream)),
                   "kr_point_states_dev")
    _lib.raise_for_flags(int(flags.item()))
    return y


def image_plane_states(cam: Camera, st: Spacetime, alpha: ArrayLike, beta: ArrayLike, *,
                       backend: Any = "hip") -> NDArray[np.float64]:
    """Past-directed null states (N, 8) for arbitrary image points of ``cam`` (benchmarks/cpu.py:140-175).

    Built on the device with :func:`initial_states`' construction; the
    reference's own construction agrees with it to 1e-12.
    """
    return image_plane_states_device(cam, st, np.asarray(alpha, np.float64), np.asarray(beta, np.float64),
                                     backend=backend).cpu().numpy()


def zamo_tetrad(st: Spacetime, r: ArrayLike, theta: ArrayLike) -> NDArray[np.float64]:
    """ZAMO tetrad rows ``e_(a)^mu``, shape ``(..., 4, 4)`` (camera.py:225-262).

    Each point is evaluated by the library's host code (``kr_zamo_frame``),
    the values the ZAMO kernels use. Raises ``ValueError`` on or inside the
    outer horizon, on the axis, or for non-finite points.
    """
    from kerrray_amd import _lib

    r_, th = np.broadcast_arrays(np.asarray(r, dtype=np.float64), np.asarray(theta, dtype=np.float64))
    if not (np.all(np.isfinite(r_)) and np.all(np.isfinite(th))):
        raise ValueError("r and theta must be finite")
    out = np.empty(r_.shape + (4, 4))
    g = np.empty(16)
    gi = np.empty(16)
    dp = ctypes.POINTER(ctypes.c_double)
    for k in np.ndindex(r_.shape):
        img = _lib.KrImage(st.mass, st.spin, float(r_[k]), float(th[k]), 1.0, 2)
        e = np.empty(16)
        _lib.check(_lib.lib().kr_zamo_frame(ctypes.byref(img), e.ctypes.data_as(dp), g.ctypes.data_as(dp),
                                            gi.ctypes.data_as(dp)), "kr_zamo_frame")
        out[k] = e.reshape(4, 4)
    return out
