"""Critical impact parameter of equatorial photons by bracketed classification.

Equatorial photons launched inward from ``r0`` with impact parameter
``b = |L_z| / E`` fall in below the critical value ``b_c`` of the circular
photon orbit and escape above it (Bardeen, Press and Teukolsky 1972). A
search holds a bracket ``[b_lo, b_hi]``, captured at ``b_lo`` and escaped at
``b_hi``, and narrows it by classifying ``n_probe`` evenly spaced interior
photons per round until it is no wider than ``tol`` or has taken
``max_iterations`` rounds. The contract (result fields, argument checks,
error messages, the optional relabelling of horizon crossings) is that of
``bisect_critical_impacts`` and ``probe_states`` in
src/kerrray/validation/kerr_bisection.py:59-300.

On the hip backend the whole search runs on the device
(``kr_impact_search``): the brackets of all searches stay in device memory,
one launch traces every open search's probes of a round, a second narrows
the brackets, and the host waits only once all rounds are enqueued. With any
other backend (the CPU oracle in the tests) the same rounds run on the
host: the brackets are arrays over the searches, each round's probes are one
``integrate_batch`` call, and the narrowing is vectorised. Both give the
reference's brackets: the probe impact parameters, the updates
``lo = max(lo, largest captured)``, ``hi = min(hi, smallest escaped)`` and the
stopping rule are the reference's, operation for operation.
"""

from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass
from typing import Any

import numpy as np
from numpy.typing import NDArray

from kerrray_amd import _lib
from kerrray_amd.backends import HipBackend, get_backend, kernel_params
from kerrray_amd.geodesics import IntegratorOptions, TerminationOptions, photon_from_constants
from kerrray_amd.geometry import Spacetime, outer_horizon
from kerrray_amd.photons import TerminationState

__all__ = ["CriticalImpactBisection", "bisect_critical_impacts", "probe_states"]

_CAPTURED = int(TerminationState.CAPTURED)
_ESCAPED = int(TerminationState.ESCAPED)


@dataclass(frozen=True)
class CriticalImpactBisection:
    """Outcome of one search (the reference's result record, kerr_bisection.py:59-88)."""

    spin: float
    prograde: bool
    b_c: float
    b_lo: float
    b_hi: float
    width: float
    iterations: int
    n_rays: int
    max_steps: int
    runtime_s: float
    horizon_crossings: int = 0


def _lz_signs(st: Spacetime, prograde: NDArray[np.bool_]) -> NDArray[np.float64]:
    """+1 where a prograde photon (a L_z > 0; L_z > 0 for a = 0) is wanted, -1 otherwise."""
    return np.where(np.asarray(prograde, bool) == (st.a >= 0.0), 1.0, -1.0)


def probe_states(st: Spacetime, bs: NDArray[np.float64], prograde: NDArray[np.bool_], r0: float,
                 integ: IntegratorOptions, term: TerminationOptions, *, horizon_crossing_as_capture: bool = False,
                 backend: Any = "hip") -> tuple[NDArray[np.int64], int, int]:
    """Trace inward equatorial photons with impact parameters ``bs`` on ``backend`` in one batch.

    Returns their termination codes, the largest accepted-step count and how
    many ``OUT_OF_DOMAIN`` photons with a finite final radius inside the
    capture margin were counted as ``CAPTURED`` (only with
    ``horizon_crossing_as_capture``).
    """
    lz = _lz_signs(st, prograde) * np.asarray(bs, np.float64)
    y0 = np.array([photon_from_constants(st, r0, 0.5 * math.pi, 1.0, float(v), 0.0, sign_r=-1, sign_theta=1)
                   for v in lz])
    engine = get_backend(backend) if isinstance(backend, str) else backend
    res = engine.integrate_batch(st, y0, integ, term)
    codes = np.asarray(res.state, np.int64).copy()
    relabelled = 0
    if horizon_crossing_as_capture:
        r_end = np.asarray(res.Y[:, 1], np.float64)
        inside = (codes == int(TerminationState.OUT_OF_DOMAIN)) & np.isfinite(r_end)
        inside &= r_end <= outer_horizon(st) + term.horizon_epsilon
        codes[inside] = _CAPTURED
        relabelled = int(inside.sum())
    return codes, int(res.n_steps.max()), relabelled


def _state_names(codes) -> list[str]:
    return [TerminationState(int(c)).name for c in codes]


def _fail_endpoints(b_lo: float, b_hi: float, prograde: bool, lo_code: int, hi_code: int) -> RuntimeError:
    lo_name, hi_name = _state_names([lo_code, hi_code])
    return RuntimeError(f"bracket [{b_lo}, {b_hi}] invalid for prograde={prograde}: "
                        f"{lo_name} at b_lo, {hi_name} at b_hi")


def _fail_states(bs, codes) -> RuntimeError:
    bad = [(float(b), name) for b, c, name in zip(bs, codes, _state_names(codes)) if c not in (_CAPTURED, _ESCAPED)]
    return RuntimeError(f"rays ended neither CAPTURED nor ESCAPED (b, state): {bad}")


def _fail_order(prograde: bool, cap_max: float, esc_min: float) -> RuntimeError:
    return RuntimeError(f"non-monotone classification in b for prograde={prograde}: captured up to "
                        f"{cap_max!r} but escaped from {esc_min!r}")


def _narrow(b: NDArray[np.float64], codes: NDArray[np.int64], prograde: NDArray[np.bool_],
            lo: NDArray[np.float64], hi: NDArray[np.float64]) -> None:
    """One round's bracket update for the searches whose probes are the rows of ``b`` (in place)."""
    if np.any((codes != _CAPTURED) & (codes != _ESCAPED)):
        raise _fail_states(b.ravel(), codes.ravel())
    cap, esc = codes == _CAPTURED, codes == _ESCAPED
    cap_max = np.where(cap, b, -np.inf).max(axis=1)
    esc_min = np.where(esc, b, np.inf).min(axis=1)
    for k in np.flatnonzero(cap.any(axis=1) & esc.any(axis=1) & (cap_max > esc_min)):
        raise _fail_order(bool(prograde[k]), float(cap_max[k]), float(esc_min[k]))
    lo[:] = np.where(cap.any(axis=1) & (cap_max > lo), cap_max, lo)
    hi[:] = np.where(esc.any(axis=1) & (esc_min < hi), esc_min, hi)


def _host_search(st, integ, term, r0, b_lo, b_hi, tol, max_iterations, n_probe, prograde, relabel, engine):
    n = prograde.size
    lo, hi = np.full(n, b_lo), np.full(n, b_hi)
    rounds = np.zeros(n, np.int64)
    kw = dict(horizon_crossing_as_capture=relabel, backend=engine)
    codes, max_steps, crossings = probe_states(st, np.tile([b_lo, b_hi], n), np.repeat(prograde, 2), r0, integ,
                                               term, **kw)
    rays = 2 * n
    for k in range(n):
        if (codes[2 * k], codes[2 * k + 1]) != (_CAPTURED, _ESCAPED):
            raise _fail_endpoints(b_lo, b_hi, bool(prograde[k]), codes[2 * k], codes[2 * k + 1])
    fractions = np.arange(1, n_probe + 1) / (n_probe + 1)
    while True:
        live = np.flatnonzero((hi - lo > tol) & (rounds < max_iterations))
        if live.size == 0:
            break
        b = lo[live, None] + (hi[live] - lo[live])[:, None] * fractions[None, :]
        codes, steps, crossed = probe_states(st, b.ravel(), np.repeat(prograde[live], n_probe), r0, integ, term,
                                             **kw)
        rays += b.size
        crossings += crossed
        max_steps = max(max_steps, steps)
        l, h = lo[live], hi[live]
        _narrow(b, codes.reshape(b.shape), prograde[live], l, h)
        lo[live], hi[live] = l, h
        rounds[live] += 1
    return lo, hi, rounds, rays, max_steps, crossings


def _device_search(st, integ, term, r0, b_lo, b_hi, tol, max_iterations, n_probe, prograde, relabel,
                   engine: HipBackend):
    n = prograde.size
    theta = 0.5 * math.pi
    o = _lib.KrImpactOpts()
    o.r0, o.theta = float(r0), theta
    o.cos2_theta, o.sin2_theta = math.cos(theta) ** 2, math.sin(theta) ** 2
    o.tol, o.capture_radius = float(tol), outer_horizon(st) + term.horizon_epsilon
    o.max_iterations, o.n_probe, o.horizon_crossing_as_capture = int(max_iterations), int(n_probe), int(relabel)
    signs = np.ascontiguousarray(_lz_signs(st, prograde))
    lo, hi = np.full(n, float(b_lo)), np.full(n, float(b_hi))
    rounds = np.zeros(n, np.int64)
    per = max(2, int(n_probe))
    probe_b, probe_code = np.zeros(n * per), np.zeros(n * per, np.int64)
    was_open = np.zeros(n, np.uint8)
    res = _lib.KrImpactResult()
    params = kernel_params(st, integ, term)
    ms = ctypes.c_double(0.0)
    with engine.handle._lock:
        _lib.check(_lib.lib().kr_impact_search(
            engine.handle.ptr, ctypes.byref(params), ctypes.byref(o), n, signs.ctypes.data, lo.ctypes.data,
            hi.ctypes.data, rounds.ctypes.data, probe_b.ctypes.data, probe_code.ctypes.data, was_open.ctypes.data,
            ctypes.byref(res), ctypes.byref(ms)), "kr_impact_search")
    engine.last_kernel_ms = ms.value
    if res.error == 1:
        for k in range(n):
            if (probe_code[2 * k], probe_code[2 * k + 1]) != (_CAPTURED, _ESCAPED):
                raise _fail_endpoints(b_lo, b_hi, bool(prograde[k]), probe_code[2 * k], probe_code[2 * k + 1])
    if res.error in (2, 3):  # re-run the failing round's checks on its probes to raise the reference's message
        live = np.flatnonzero(was_open)
        b = probe_b.reshape(n, per)[live, :n_probe]
        codes = probe_code.reshape(n, per)[live, :n_probe]
        _narrow(b, codes, prograde[live], lo[live].copy(), hi[live].copy())
        raise RuntimeError(f"the device search reported error {res.error} that the host check did not reproduce")
    return lo, hi, rounds, int(res.rays), int(res.max_steps), int(res.crossings)


def bisect_critical_impacts(st: Spacetime, integ: IntegratorOptions, term: TerminationOptions, *, r0: float,
                            b_lo: float, b_hi: float, tol: float, max_iterations: int, n_probe: int = 8,
                            directions: tuple[bool, ...] = (True, False), horizon_crossing_as_capture: bool = False,
                            backend: Any = "hip") -> tuple[CriticalImpactBisection, ...]:
    """Bracketing search for ``b_c`` in each of ``directions`` (``True`` prograde, ``False`` retrograde).

    ``ValueError`` for an invalid bracket, tolerance or probe count;
    ``RuntimeError`` for a photon that ends neither ``CAPTURED`` nor
    ``ESCAPED``, misclassified bracket endpoints, or a captured probe above an
    escaped one.
    """
    if not (0.0 <= b_lo < b_hi):
        raise ValueError(f"need 0 <= b_lo < b_hi, got {b_lo!r}, {b_hi!r}")
    if not tol > 0.0:
        raise ValueError(f"tol must be > 0, got {tol!r}")
    if n_probe < 1 or max_iterations < 1:
        raise ValueError("n_probe and max_iterations must be >= 1")
    t0 = time.perf_counter()
    prograde = np.array(directions, dtype=bool)
    engine = get_backend(backend) if isinstance(backend, str) else backend
    search = _device_search if isinstance(engine, HipBackend) else _host_search
    lo, hi, rounds, rays, max_steps, crossings = search(st, integ, term, r0, float(b_lo), float(b_hi), tol,
                                                        max_iterations, n_probe, prograde,
                                                        horizon_crossing_as_capture, engine)
    runtime = time.perf_counter() - t0
    return tuple(
        CriticalImpactBisection(spin=st.spin, prograde=bool(p), b_c=0.5 * (float(l) + float(h)), b_lo=float(l),
                                b_hi=float(h), width=float(h) - float(l), iterations=int(r), n_rays=rays,
                                max_steps=max_steps, runtime_s=runtime, horizon_crossings=crossings)
        for p, l, h, r in zip(prograde, lo, hi, rounds)
    )
