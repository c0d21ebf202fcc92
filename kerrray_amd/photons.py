"""Termination codes and the batch result (host side).

Same codes and fields as the reference (src/kerrray/photons/classification.py:60-69,
src/kerrray/photons/trajectories.py:100-138).
"""

from __future__ import annotations

from dataclasses import dataclass
from enum import IntEnum

import numpy as np
from numpy.typing import NDArray

__all__ = ["AXIS_SIN_TOLERANCE", "FAILURE_STATES", "THETA_DOMAIN_TOLERANCE", "BatchResult", "TerminationState"]

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
