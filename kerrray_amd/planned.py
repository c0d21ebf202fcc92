"""Planned image traces: longest-first schedule, and the FP32 bulk with FP64 repair.

Both run the same four device kernels (csrc/kr_mixed.hip): a plan that
orders the pixels edge band -> inside -> outside using Bardeen's curve, a
persistent trace of that work list, a check of the traced mask, and an FP64
repair of every pixel near the capture boundary (or failed) that was not
already traced in FP64.

* ``fp32_bulk=False``: every photon in FP64, identical bit for bit to
  :func:`kerrray_amd.shadow.compute_shadow`; only the schedule differs.
* ``fp32_bulk=True``: photons farther than ``band_px`` pixels from the
  boundary take FP32 right-hand sides (FP64 state, update, error norm and
  step control; optionally FP64 attempts near the horizon, the axis and
  strong-field radial turning points). The mask is then that of the FP64
  trace: every pixel whose class could differ is re-traced in FP64.
"""

from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, field

import numpy as np

from kerrray_amd import _lib
from kerrray_amd.backends import HipBackend, HostBatch, get_backend, kernel_params
from kerrray_amd.boundary import finite_distance_scale, polygon_centroid, polygon_radii
from kerrray_amd.camera import BardeenCamera, Camera, image_struct
from kerrray_amd.geodesics import IntegratorOptions, TerminationOptions
from kerrray_amd.geometry import Spacetime
from kerrray_amd.orbits import shadow_curve
from kerrray_amd.photons import BatchResult, TerminationState
from kerrray_amd.shadow import ShadowImage, state_counts

__all__ = ["EdgeTable", "PlannedOptions", "compute_shadow_planned", "edge_table", "trace_planned"]


@dataclass(frozen=True)
class PlannedOptions:
    fp32_bulk: bool = True
    band_px: int = 2
    # FP64 attempts for r < r_+ + horizon_margin, |sin theta| < axis_sin, or
    # r < turn_rmax with |dr/dlambda| < turn_drdl.  Off by default: the mask
    # comes from the FP64 edge band and repair either way, and on the test
    # images these switches change neither the mask nor the median or 99th
    # percentile error of the escape directions, while they cost a fifth of
    # the time (wavefronts whose lanes disagree run both precisions).
    horizon_margin: float = 0.0  # M
    axis_sin: float = 0.0
    turn_rmax: float = 10.0  # M
    turn_drdl: float = 0.0
    refill_interval: int = 8
    block_size: int = 128
    longest_first: bool = True
    band_priority: bool = False
    # 0: automatic (one wavefront per SIMD for the split FP32 form up to 2^19
    # pixels, where its longest FP64 photons bound the time; two otherwise).
    waves_per_simd: int = 0
    # Photons with |L_z| = |alpha sin theta_o| < axis_lz (M) pass close to a
    # pole, where the FP32 error estimate can stall the step size: they are
    # planned in FP64 from the start, with the edge band.
    axis_lz: float = 0.1
    # FP32 bulk only: wavefronts trace either the FP64 band and strip or FP32
    # photons, never both in one wavefront (1, or True).  At one wavefront per
    # SIMD and rtol >= 1e-7 each FP64 photon is then carried by two lanes
    # that share out its trigonometry (the same rows, a shorter longest
    # photon); 2: always paired, 3: never.
    split_roles: int = 1
    # Share of the resident-wavefront slots the trace kernel is launched on,
    # in (0, 1]; 0: automatic (7/8 for the split FP32 form at one wavefront
    # per SIMD up to 2^18 pixels and rtol >= 1e-6, 15/16 when its FP64
    # photons are paired, whose FP32 photons finish
    # well before the longest FP64 photon, so that fewer active wavefronts
    # leave the clock higher for it; all slots otherwise).
    grid_fraction: float = 0.0

    def kr(self) -> _lib.KrPlannedOpts:
        o = _lib.KrPlannedOpts()
        o.horizon_margin = self.horizon_margin
        o.axis_sin = self.axis_sin
        o.turn_rmax = self.turn_rmax
        o.turn_drdl = self.turn_drdl
        o.band_px = self.band_px
        o.fp32_bulk = 1 if self.fp32_bulk else 0
        o.refill_interval = self.refill_interval
        o.block_size = self.block_size
        o.longest_first = 1 if self.longest_first else 0
        o.band_priority = 1 if self.band_priority else 0
        o.waves_per_simd = self.waves_per_simd
        o.axis_lz = self.axis_lz
        o.split_roles = int(self.split_roles)
        o.grid_fraction = self.grid_fraction
        return o


@dataclass(frozen=True)
class EdgeTable:
    """Polar radii of Bardeen's curve about its centroid, at angles 2 pi k / n."""

    centre: tuple[float, float]
    radii: np.ndarray = field(repr=False)

    @property
    def n_angles(self) -> int:
        return int(self.radii.size)


_EDGE_CACHE: dict[tuple, EdgeTable] = {}


def edge_table(st: Spacetime, inclination_deg: float, n_angles: int = 2048, curve_points: int = 4000) -> EdgeTable:
    key = (st.mass, st.spin, float(inclination_deg), n_angles, curve_points)
    if key not in _EDGE_CACHE:
        a, b = shadow_curve(st, inclination_deg, curve_points)
        c = polygon_centroid(a, b)
        ang = 2.0 * np.pi * np.arange(n_angles) / n_angles
        radii = np.concatenate([polygon_radii(a, b, c, ang[k:k + 256]) for k in range(0, n_angles, 256)])
        _EDGE_CACHE[key] = EdgeTable(centre=c, radii=np.ascontiguousarray(radii))
    return _EDGE_CACHE[key]


@dataclass
class PlannedStats:
    planned_band: int
    planned_inside: int
    planned_outside: int
    repaired: int
    fp64_attempts: int
    total_attempts: int
    kernel_ms: float

    @property
    def fp64_fraction(self) -> float:
        return self.fp64_attempts / max(1, self.total_attempts)


def trace_planned(engine: HipBackend, st: Spacetime, cam: Camera | BardeenCamera, integ: IntegratorOptions,
                  term: TerminationOptions, options: PlannedOptions = PlannedOptions(),
                  ) -> tuple[BatchResult, _lib.KrPlannedStats, float]:
    """Every pixel of ``cam`` through ``kr_trace_image_planned``: the rows, the plan's counts and the call's
    device time (ms, also left in ``engine.last_kernel_ms``)."""
    t0 = time.perf_counter()
    table = edge_table(st, cam.inclination_deg)
    if isinstance(cam, Camera):
        # The ZAMO camera sees Bardeen's curve shrunk by the finite-distance
        # factor; the plan only needs it to a fraction of a pixel.
        k = finite_distance_scale(st, cam.radius)
        table = EdgeTable(centre=(table.centre[0] / k, table.centre[1] / k), radii=table.radii / k)
    params = kernel_params(st, integ, term)
    img = image_struct(st, cam)
    out = HostBatch(cam.n_rays, np.dtype(np.float64), False)
    stats = _lib.KrPlannedStats()
    ms = ctypes.c_double(0.0)
    opts = options.kr()
    with engine.handle._lock:
        _lib.check(
            _lib.lib().kr_trace_image_planned(
                engine.handle.ptr, ctypes.byref(params), ctypes.byref(img), ctypes.byref(opts), table.centre[0],
                table.centre[1], table.radii.ctypes.data, table.n_angles, ctypes.byref(out.outputs),
                ctypes.byref(stats), ctypes.byref(ms)),
            "kr_trace_image_planned",
        )
    engine.last_kernel_ms = ms.value
    return out.result(time.perf_counter() - t0), stats, ms.value


def compute_shadow_planned(
    st: Spacetime,
    cam: Camera | BardeenCamera,
    integ: IntegratorOptions,
    term: TerminationOptions,
    *,
    backend: str | HipBackend = "hip",
    options: PlannedOptions = PlannedOptions(),
) -> tuple[ShadowImage, PlannedStats]:
    engine = get_backend(backend) if isinstance(backend, str) else backend
    if not isinstance(engine, HipBackend):
        raise TypeError("planned traces run on the hip backend")
    result, stats, ms = trace_planned(engine, st, cam, integ, term, options)
    alpha, beta = cam.pixel_coordinates()
    state = result.state.reshape(cam.shape)
    image = ShadowImage(alpha=alpha, beta=beta, state=state, captured=state == int(TerminationState.CAPTURED),
                        counts=state_counts(state), result=result, camera=cam, spacetime=st)
    ps = PlannedStats(stats.planned_band, stats.planned_inside, stats.planned_outside, stats.repaired,
                      stats.fp64_attempts, int(np.sum(result.n_steps + result.n_rejected)), ms)
    return image, ps
