"""Termination codes and the batch result (host side).

Same codes and fields as the reference (src/kerrray/photons/classification.py:60-69,
src/kerrray/photons/trajectories.py:100-138).
"""

from __future__ import annotations

from dataclasses import dataclass
from enum import IntEnum

import numpy as np
from numpy.typing import NDArray

__all__ = ["AXIS_SIN_TOLERANCE", "FAILURE_STATES", "THETA_DOMAIN_TOLERANCE", "BatchResult", "ResultRows",
           "TerminationState"]

THETA_DOMAIN_TOLERANCE = 1e-9
AXIS_SIN_TOLERANCE = 1e-12


class TerminationState(IntEnum):
    RUNNING = 0
    ESCAPED = 1
    CAPTURED = 2
    MAX_AFFINE_PARAMETER = 3
    NUMERICAL_FAILURE = 4
    OUT_OF_DOMAIN = 5
    DISK_HIT = 6


FAILURE_STATES = (TerminationState.NUMERICAL_FAILURE, TerminationState.OUT_OF_DOMAIN)


@dataclass
class BatchResult:
    """Final states and per-photon summaries of one batch."""

    Y: NDArray[np.floating]
    state: NDArray[np.int64]
    n_steps: NDArray[np.int64]
    lam: NDArray[np.float64]
    max_null_error: NDArray[np.float64]
    max_energy_drift: NDArray[np.float64]
    max_lz_drift: NDArray[np.float64]
    max_carter_drift: NDArray[np.float64]
    runtime_s: float
    event_Y: NDArray[np.floating] | None = None
    event_hit: NDArray[np.bool_] | None = None
    n_rejected: NDArray[np.int64] | None = None

    @property
    def n_rays(self) -> int:
        return int(self.Y.shape[0])

    def count(self, state: TerminationState) -> int:
        return int(np.count_nonzero(self.state == int(state)))


class ResultRows:
    """Host arrays for the per-photon results of ``n`` photons, filled in place.

    A trace of many photons is written into one allocation: each chunk of a
    chunked trace goes into its own row range (``rows(lo, hi)`` is a view on
    the same memory), so the batch is never concatenated. ``event_hit`` is
    held as bytes (what the device writes) and handed out as ``bool``.
    """

    def __init__(self, n: int, y_dtype, events: bool, rejected: bool = True, *, _views: dict | None = None) -> None:
        if _views is not None:
            self.__dict__.update(_views)
            return
        self.n = int(n)
        self.Y = np.zeros((n, 8), dtype=y_dtype)
        self.lam = np.zeros(n)
        self.n_steps = np.zeros(n, np.int64)
        self.n_rejected = np.zeros(n, np.int64) if rejected else None
        self.state = np.zeros(n, np.int64)
        self.max_null_error = np.zeros(n)
        self.max_energy_drift = np.zeros(n)
        self.max_lz_drift = np.zeros(n)
        self.max_carter_drift = np.zeros(n)
        self.event_Y = np.zeros((n, 8), dtype=y_dtype) if events else None
        self.event_hit = np.zeros(n, np.uint8) if events else None

    _ARRAYS = ("Y", "lam", "n_steps", "n_rejected", "state", "max_null_error", "max_energy_drift", "max_lz_drift",
               "max_carter_drift", "event_Y", "event_hit")

    def rows(self, lo: int, hi: int) -> ResultRows:
        """Rows ``[lo, hi)`` as a store of their own that writes through to this one."""
        views = {name: (None if getattr(self, name) is None else getattr(self, name)[lo:hi]) for name in self._ARRAYS}
        views["n"] = hi - lo
        return ResultRows(0, None, False, _views=views)

    def fill(self, lo: int, part: BatchResult) -> None:
        """Copy a chunk's results into rows ``[lo, lo + part.n_rays)``.

        An optional field the chunk lacks (events, rejected counts) is dropped
        from the whole store, as a concatenation would drop it.
        """
        hi = lo + part.n_rays
        for name in self._ARRAYS:
            dst = getattr(self, name)
            if dst is None:
                continue
            src = getattr(part, name)
            if src is None:
                setattr(self, name, None)
                continue
            dst[lo:hi] = src

    def result(self, runtime_s: float) -> BatchResult:
        return BatchResult(
            Y=self.Y, state=self.state, n_steps=self.n_steps, lam=self.lam, max_null_error=self.max_null_error,
            max_energy_drift=self.max_energy_drift, max_lz_drift=self.max_lz_drift,
            max_carter_drift=self.max_carter_drift, runtime_s=float(runtime_s), event_Y=self.event_Y,
            event_hit=None if self.event_hit is None else self.event_hit.astype(bool), n_rejected=self.n_rejected,
        )
