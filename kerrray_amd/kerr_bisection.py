"""Critical impact parameter by bracketed classification search, on any backend.

Mirrors ``CriticalImpactBisection``, ``probe_states`` and
``bisect_critical_impacts`` of the reference
(src/kerrray/validation/kerr_bisection.py:59-300). Equatorial photons launched
inward from ``r0`` with impact parameter ``b = |L_z| / E`` are captured below
the critical value ``b_c`` of the circular photon orbit and escape above it
(Bardeen, Press and Teukolsky 1972). The search keeps a bracket
``[b_lo, b_hi]``, captured at ``b_lo`` and escaped at ``b_hi``, and classifies
``n_probe`` interior rays per iteration. Prograde and retrograde searches
share one batch.

The reference integrates the probes with its NumPy integrator directly;
here ``probe_states`` takes a ``backend`` (the ``hip`` backend by default),
whose ``integrate_batch`` traces the batch. Initial states come from
:func:`kerrray_amd.geodesics.photon_from_constants`, the reference's
expressions. Any ray that ends neither ``CAPTURED`` nor ``ESCAPED``, an
invalid starting bracket, or an ``ESCAPED`` probe below a ``CAPTURED`` one
raises ``RuntimeError``, as in the reference.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass
from typing import Any

import numpy as np
from numpy.typing import NDArray

from kerrray_amd.backends import get_backend
from kerrray_amd.geodesics import IntegratorOptions, TerminationOptions, photon_from_constants
from kerrray_amd.geometry import Spacetime, outer_horizon
from kerrray_amd.photons import TerminationState

__all__ = ["CriticalImpactBisection", "bisect_critical_impacts", "probe_states"]


@dataclass(frozen=True)
class CriticalImpactBisection:
    """Result of one bracketing search (kerr_bisection.py:59-88)."""

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


def _lz_sign(st: Spacetime, prograde: bool) -> float:
    """``+1`` or ``-1`` such that ``a L_z > 0`` for prograde (``L_z > 0`` for ``a = 0``)."""
    return 1.0 if prograde == (st.a >= 0.0) else -1.0


def probe_states(st: Spacetime, bs: NDArray[np.float64], prograde: NDArray[np.bool_], r0: float,
                 integ: IntegratorOptions, term: TerminationOptions, *, horizon_crossing_as_capture: bool = False,
                 backend: Any = "hip") -> tuple[NDArray[np.int64], int, int]:
    """Classify inward equatorial rays with impact parameters ``bs`` in one batch (kerr_bisection.py:96-127).

    Returns the termination codes, the largest accepted-step count and the
    number of rays relabelled ``CAPTURED`` (an ``OUT_OF_DOMAIN`` ray whose
    finite final radius is within the capture margin, when
    ``horizon_crossing_as_capture`` is set).
    """
    y0 = np.stack([
        photon_from_constants(st, r0, 0.5 * math.pi, 1.0, _lz_sign(st, bool(p)) * b, 0.0, sign_r=-1, sign_theta=1)
        for b, p in zip(bs, prograde, strict=True)
    ])
    engine = get_backend(backend) if isinstance(backend, str) else backend
    result = engine.integrate_batch(st, y0, integ, term)
    states = np.array(result.state, dtype=np.int64, copy=True)
    n_crossed = 0
    if horizon_crossing_as_capture:
        r_end = np.asarray(result.Y[:, 1], dtype=np.float64)
        crossed = (states == TerminationState.OUT_OF_DOMAIN) & np.isfinite(r_end) & (
            r_end <= outer_horizon(st) + term.horizon_epsilon)
        states[crossed] = TerminationState.CAPTURED
        n_crossed = int(np.count_nonzero(crossed))
    return states, int(np.max(result.n_steps)), n_crossed


class _Search:
    """The bracket of one search."""

    def __init__(self, prograde: bool, b_lo: float, b_hi: float) -> None:
        self.prograde = prograde
        self.lo = b_lo
        self.hi = b_hi
        self.iterations = 0

    @property
    def width(self) -> float:
        return self.hi - self.lo

    def update(self, bs: NDArray[np.float64], states: NDArray[np.int64]) -> None:
        captured = bs[states == TerminationState.CAPTURED]
        escaped = bs[states == TerminationState.ESCAPED]
        if captured.size and escaped.size and captured.max() > escaped.min():
            raise RuntimeError(f"non-monotone classification in b for prograde={self.prograde}: captured up to "
                               f"{captured.max()!r} but escaped from {escaped.min()!r}")
        if captured.size:
            self.lo = max(self.lo, float(captured.max()))
        if escaped.size:
            self.hi = min(self.hi, float(escaped.min()))
        self.iterations += 1


def _check_states(bs: NDArray[np.float64], states: NDArray[np.int64]) -> None:
    bad = [(float(b), TerminationState(int(s)).name) for b, s in zip(bs, states, strict=True)
           if int(s) not in (TerminationState.CAPTURED, TerminationState.ESCAPED)]
    if bad:
        raise RuntimeError(f"rays ended neither CAPTURED nor ESCAPED (b, state): {bad}")


def bisect_critical_impacts(st: Spacetime, integ: IntegratorOptions, term: TerminationOptions, *, r0: float,
                            b_lo: float, b_hi: float, tol: float, max_iterations: int, n_probe: int = 8,
                            directions: tuple[bool, ...] = (True, False), horizon_crossing_as_capture: bool = False,
                            backend: Any = "hip") -> tuple[CriticalImpactBisection, ...]:
    """Bracketing search for ``b_c`` in each of ``directions`` (prograde ``True``) (kerr_bisection.py:177-300).

    Raises ``ValueError`` for an invalid bracket, tolerance or probe count,
    and ``RuntimeError`` for a ray that ends neither ``CAPTURED`` nor
    ``ESCAPED``, misclassified bracket endpoints, or a non-monotone outcome.
    """
    if not (0.0 <= b_lo < b_hi):
        raise ValueError(f"need 0 <= b_lo < b_hi, got {b_lo!r}, {b_hi!r}")
    if not tol > 0.0:
        raise ValueError(f"tol must be > 0, got {tol!r}")
    if n_probe < 1 or max_iterations < 1:
        raise ValueError("n_probe and max_iterations must be >= 1")
    t0 = time.perf_counter()
    searches = [_Search(p, b_lo, b_hi) for p in directions]
    probe = dict(horizon_crossing_as_capture=horizon_crossing_as_capture, backend=backend)
    bs = np.array([b_lo, b_hi] * len(searches))
    pro = np.array([p for s in searches for p in (s.prograde, s.prograde)])
    states, max_steps, n_crossed = probe_states(st, bs, pro, r0, integ, term, **probe)
    n_rays = bs.size
    for k, s in enumerate(searches):
        lo_state, hi_state = TerminationState(int(states[2 * k])), TerminationState(int(states[2 * k + 1]))
        if lo_state != TerminationState.CAPTURED or hi_state != TerminationState.ESCAPED:
            raise RuntimeError(f"bracket [{b_lo}, {b_hi}] invalid for prograde={s.prograde}: "
                               f"{lo_state.name} at b_lo, {hi_state.name} at b_hi")
    fractions = np.arange(1, n_probe + 1) / (n_probe + 1)
    while True:
        open_searches = [s for s in searches if s.width > tol and s.iterations < max_iterations]
        if not open_searches:
            break
        bs = np.concatenate([s.lo + s.width * fractions for s in open_searches])
        pro = np.concatenate([np.full(n_probe, s.prograde) for s in open_searches])
        states, steps, crossed = probe_states(st, bs, pro, r0, integ, term, **probe)
        _check_states(bs, states)
        n_rays += bs.size
        n_crossed += crossed
        max_steps = max(max_steps, steps)
        for k, s in enumerate(open_searches):
            sl = slice(k * n_probe, (k + 1) * n_probe)
            s.update(bs[sl], states[sl])
    runtime = time.perf_counter() - t0
    return tuple(
        CriticalImpactBisection(spin=st.spin, prograde=s.prograde, b_c=0.5 * (s.lo + s.hi), b_lo=s.lo, b_hi=s.hi,
                                width=s.width, iterations=s.iterations, n_rays=n_rays, max_steps=max_steps,
                                runtime_s=runtime, horizon_crossings=n_crossed)
        for s in searches
    )
