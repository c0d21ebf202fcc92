"""Chunked tracing of initial states on a backend.

Same contract as the reference's ``trace_rays`` and ``concatenate_results``
(src/kerrray/raytracing/rays.py:31-117): results are in input order and do
not depend on the chunking.
"""

from __future__ import annotations

import time
from collections.abc import Callable

import numpy as np
from numpy.typing import ArrayLike, NDArray

from kerrray_amd.backends import Backend, get_backend
from kerrray_amd.geodesics import STATE_SIZE, EventOptions, IntegratorOptions, TerminationOptions
from kerrray_amd.geometry import Spacetime
from kerrray_amd.photons import BatchResult

__all__ = ["DEFAULT_CHUNK_SIZE", "ProgressCallback", "concatenate_results", "trace_rays"]

DEFAULT_CHUNK_SIZE = 1 << 22
"""Rays per chunk. A GPU batch is cheapest when it is large: 4 Mi photons keep
the device busy while their inputs and outputs (~0.5 GB) fit comfortably."""

ProgressCallback = Callable[[float], None]


def _optional_concat(parts: list[NDArray | None]) -> NDArray | None:
    if any(part is None for part in parts):
        return None
    return np.concatenate([part for part in parts if part is not None], axis=0)


def concatenate_results(parts: list[BatchResult], runtime_s: float | None = None) -> BatchResult:
    """Join per-ray fields of ``parts`` in order (rays.py:46-69)."""
    if not parts:
        raise ValueError("concatenate_results needs at least one BatchResult")
    if len(parts) == 1 and runtime_s is None:
        return parts[0]
    return BatchResult(
        Y=np.concatenate([p.Y for p in parts], axis=0),
        state=np.concatenate([p.state for p in parts]),
        n_steps=np.concatenate([p.n_steps for p in parts]),
        lam=np.concatenate([p.lam for p in parts]),
        max_null_error=np.concatenate([p.max_null_error for p in parts]),
        max_energy_drift=np.concatenate([p.max_energy_drift for p in parts]),
        max_lz_drift=np.concatenate([p.max_lz_drift for p in parts]),
        max_carter_drift=np.concatenate([p.max_carter_drift for p in parts]),
        runtime_s=float(sum(p.runtime_s for p in parts)) if runtime_s is None else float(runtime_s),
        event_Y=_optional_concat([p.event_Y for p in parts]),
        event_hit=_optional_concat([p.event_hit for p in parts]),
        n_rejected=_optional_concat([p.n_rejected for p in parts]),
    )


def trace_rays(
    st: Spacetime,
    Y0: ArrayLike,
    integ: IntegratorOptions,
    term: TerminationOptions,
    *,
    backend: str | Backend = "hip",
    events: EventOptions | None = None,
    progress: ProgressCallback | None = None,
    chunk_size: int = DEFAULT_CHUNK_SIZE,
) -> BatchResult:
    """Integrate the states ``Y0`` (N, 8) in chunks on ``backend`` (rays.py:72-117)."""
    y0 = np.asarray(Y0)
    if y0.ndim != 2 or y0.shape[1] != STATE_SIZE:
        raise ValueError(f"Y0 must have shape (N, {STATE_SIZE}), got {y0.shape}")
    n_rays = y0.shape[0]
    if n_rays < 1:
        raise ValueError("Y0 must contain at least one ray")
    if chunk_size < 1:
        raise ValueError(f"chunk_size must be >= 1, got {chunk_size!r}")
    engine = get_backend(backend) if isinstance(backend, str) else backend
    t0 = time.perf_counter()
    parts: list[BatchResult] = []
    for start in range(0, n_rays, chunk_size):
        stop = min(start + chunk_size, n_rays)
        parts.append(engine.integrate_batch(st, y0[start:stop], integ, term, events=events))
        if progress is not None:
            progress(stop / n_rays)
    return concatenate_results(parts, runtime_s=time.perf_counter() - t0)
