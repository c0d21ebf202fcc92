"""Shadow-edge location by bisection along image-plane directions, on the GPU.

``bisect_edge`` keeps the signature and semantics of the reference's study
(scripts/study_edge_bisection.py:53-68): for each direction phi_k about
``centre`` the bracket [lo, hi] (captured side at lo, escaping side at hi)
is halved ``iterations`` times, mid = (lo + hi) / 2, escaped -> hi = mid,
otherwise lo = mid; the integrator is DP5(4) with atol = 1e-2 rtol,
max_steps = 2e6, lambda_max = 4 r_o, and termination at r_esc = r_o with
horizon margin 1e-6. Photons start from Bardeen image-plane initial
conditions at the observer. The whole loop runs on the device
(``kr_bisect_edge``), one lane per (direction, probe).

``n_probe > 1`` traces that many evenly spaced interior points per round
and keeps the sub-bracket around the first escaping one, shrinking the
bracket by n_probe + 1 per round instead of 2 (the multi-probe idea of
src/kerrray/validation/kerr_bisection.py:12-18): fewer, wider launches for
the same final bracket.

``shadow_edge`` is the image-level driver: directions about the centroid of
Bardeen's curve, a +-2 % bracket around the curve, and the edge error
against the curve (for Bardeen initial conditions the captured region is the
inside of the curve itself, so no finite-distance factor applies).
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
from numpy.typing import ArrayLike, NDArray

from kerrray_amd import _lib
from kerrray_amd.backends import Backend, HipBackend, get_backend, kernel_params
from kerrray_amd.boundary import edge_error, polygon_centroid, polygon_radii
from kerrray_amd.geodesics import IntegratorOptions, TerminationOptions
from kerrray_amd.geometry import Spacetime
from kerrray_amd.orbits import shadow_curve
from kerrray_amd.shadow import kr_image

__all__ = ["EdgeResult", "bisect_edge", "bisection_options", "rounds_for_width", "shadow_edge"]


def bisection_options(rtol: float, r_o: float) -> tuple[IntegratorOptions, TerminationOptions]:
    """The integrator and termination settings of study_edge_bisection.py:54-55."""
    integ = IntegratorOptions(method="rk45", rtol=rtol, atol=rtol * 1e-2, max_steps=2_000_000, lambda_max=4.0 * r_o)
    term = TerminationOptions(horizon_epsilon=1e-6, escape_radius=r_o)
    return integ, term


@dataclass
class EdgeResult:
    radius: NDArray[np.float64]  # (lo + hi) / 2 per direction
    width: NDArray[np.float64]  # hi - lo
    failures: int  # traced photons neither escaped nor captured
    nonmonotone: int  # (direction, round) pairs whose probes were not ordered inside -> outside
    lo: NDArray[np.float64]
    hi: NDArray[np.float64]
    kernel_ms: float


def rounds_for_width(width0: float, width: float, n_probe: int) -> int:
    """Rounds needed to shrink a bracket from width0 to <= width with n_probe probes per round."""
    return max(0, int(math.ceil(math.log(width0 / width) / math.log(n_probe + 1) - 1e-12)))


def bisect_edge(
    st: Spacetime,
    incl: float,
    r_o: float,
    rtol: float,
    centre: tuple[float, float],
    angles: ArrayLike,
    r_lo: ArrayLike,
    r_hi: ArrayLike,
    iterations: int,
    backend: str | Backend = "hip",
    *,
    n_probe: int = 1,
    integ: IntegratorOptions | None = None,
    term: TerminationOptions | None = None,
) -> EdgeResult:
    engine = get_backend(backend) if isinstance(backend, str) else backend
    if not isinstance(engine, HipBackend):
        raise TypeError("bisect_edge runs on the hip backend")
    d_integ, d_term = bisection_options(rtol, r_o)
    integ = integ or d_integ
    term = term or d_term
    phi = np.ascontiguousarray(np.asarray(angles, dtype=np.float64).ravel())
    lo = np.array(np.broadcast_to(np.asarray(r_lo, dtype=np.float64), phi.shape))
    hi = np.array(np.broadcast_to(np.asarray(r_hi, dtype=np.float64), phi.shape))
    cos_k = np.ascontiguousarray(np.cos(phi))
    sin_k = np.ascontiguousarray(np.sin(phi))
    params = kernel_params(st, integ, term)
    img = kr_image(st, r_o, incl)
    failures = ctypes.c_int64(0)
    nonmono = ctypes.c_int64(0)
    ms = ctypes.c_double(0.0)
    with engine.handle._lock:
        _lib.check(
            _lib.lib().kr_bisect_edge(
                engine.handle.ptr, ctypes.byref(params), ctypes.byref(img), float(centre[0]), float(centre[1]),
                cos_k.ctypes.data, sin_k.ctypes.data, phi.size, lo.ctypes.data, hi.ctypes.data, int(iterations),
                int(n_probe), ctypes.byref(failures), ctypes.byref(nonmono), ctypes.byref(ms),
            ),
            "kr_bisect_edge",
        )
    return EdgeResult(radius=0.5 * (lo + hi), width=hi - lo, failures=int(failures.value),
                      nonmonotone=int(nonmono.value), lo=lo, hi=hi, kernel_ms=ms.value)


def shadow_edge(
    st: Spacetime,
    incl: float,
    r_o: float,
    rtol: float,
    n_angles: int,
    *,
    backend: str | Backend = "hip",
    n_probe: int = 1,
    width: float = 1e-11,
    curve_points: int = 40000,
) -> tuple[EdgeResult, NDArray[np.float64], tuple[float, float], dict[str, float]]:
    """Edge radii along ``n_angles`` uniform directions and their error against Bardeen's curve."""
    curve = shadow_curve(st, incl, curve_points)
    centre = polygon_centroid(*curve)
    angles = 2.0 * np.pi * np.arange(n_angles, dtype=np.float64) / n_angles
    r_ana = polygon_radii(curve[0], curve[1], centre, angles)
    lo, hi = 0.98 * r_ana, 1.02 * r_ana
    rounds = rounds_for_width(float(np.max(hi - lo)), width, n_probe)
    res = bisect_edge(st, incl, r_o, rtol, centre, angles, lo, hi, rounds, backend, n_probe=n_probe)
    return res, angles, centre, edge_error(res.radius, angles, centre, curve)
