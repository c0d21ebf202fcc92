"""Tracing a batch of initial states on a backend, chunk by chunk, into one result.

The contract is that of the reference's ``trace_rays`` and
``concatenate_results`` (src/kerrray/raytracing/rays.py:31-117): the same
arguments, checks and messages; results in input order and independent of
the chunking; ``progress`` called with the completed fraction after each
chunk, ending at exactly 1.0; ``runtime_s`` the wall time of the call.

The design differs. The result arrays of all N photons are allocated once
(:class:`~kerrray_amd.photons.ResultRows`) and every chunk is written into
its own row range, so nothing is concatenated. A backend that can write into
a row store (the hip backend's ``integrate_into``) writes there directly;
any other backend's per-chunk results are copied in. On hip the default
chunk is large enough that an image is one device call.
"""

from __future__ import annotations

import time
from collections.abc import Callable, Iterator

import numpy as np
from numpy.typing import ArrayLike

from kerrray_amd.backends import Backend, get_backend
from kerrray_amd.geodesics import STATE_SIZE, EventOptions, IntegratorOptions, TerminationOptions
from kerrray_amd.geometry import Spacetime
from kerrray_amd.photons import BatchResult, ResultRows

__all__ = ["DEFAULT_CHUNK_SIZE", "ProgressCallback", "concatenate_results", "trace_rays"]

DEFAULT_CHUNK_SIZE = 1 << 22
"""Rays per chunk. A device call is cheapest when it is large: 4 Mi photons keep
the GPU busy while their inputs and outputs (~0.5 GB) fit comfortably."""

ProgressCallback = Callable[[float], None]


def _row_ranges(n: int, chunk: int) -> Iterator[tuple[int, int]]:
    lo = 0
    while lo < n:
        hi = lo + chunk if n - lo > chunk else n
        yield lo, hi
        lo = hi


def _store_like(part: BatchResult, n: int) -> ResultRows:
    """An n-row store with the fields and dtypes of ``part``."""
    return ResultRows(n, part.Y.dtype, part.event_Y is not None, part.n_rejected is not None)


def concatenate_results(parts: list[BatchResult], runtime_s: float | None = None) -> BatchResult:
    """The per-ray fields of ``parts`` joined in order; ``runtime_s`` defaults to the parts' sum."""
    if not parts:
        raise ValueError("concatenate_results needs at least one BatchResult")
    if runtime_s is None:
        if len(parts) == 1:
            return parts[0]
        runtime_s = float(sum(p.runtime_s for p in parts))
    store = _store_like(parts[0], sum(p.n_rays for p in parts))
    lo = 0
    for part in parts:
        store.fill(lo, part)
        lo += part.n_rays
    return store.result(runtime_s)


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
    """Integrate the states ``Y0`` (N, 8) on ``backend``, at most ``chunk_size`` per call."""
    states = np.asarray(Y0)
    if states.ndim != 2 or states.shape[1] != STATE_SIZE:
        raise ValueError(f"Y0 must have shape (N, {STATE_SIZE}), got {states.shape}")
    n = states.shape[0]
    if n < 1:
        raise ValueError("Y0 must contain at least one ray")
    if chunk_size < 1:
        raise ValueError(f"chunk_size must be >= 1, got {chunk_size!r}")
    engine = get_backend(backend) if isinstance(backend, str) else backend
    write_into = getattr(engine, "integrate_into", None)
    t0 = time.perf_counter()
    store: ResultRows | None = None
    if write_into is not None:
        store = ResultRows(n, integ.np_dtype, events is not None and events.disk_plane)
    for lo, hi in _row_ranges(n, chunk_size):
        if write_into is not None:
            write_into(st, states[lo:hi], integ, term, events, store.rows(lo, hi))
        else:
            part = engine.integrate_batch(st, states[lo:hi], integ, term, events=events)
            if store is None:
                store = _store_like(part, n)
            store.fill(lo, part)
        if progress is not None:
            progress(hi / n)
    return store.result(time.perf_counter() - t0)
