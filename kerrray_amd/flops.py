"""Algorithmic floating-point work of the per-photon path, in one place.

Counts are of the reference algorithm (src/kerrray/raytracing/backends/numba_kernel.py),
one flop per add, subtract, multiply, divide or square root, negation and
comparisons free, transcendental calls counted separately:

* one right-hand side: 79 in ``inverse_metric_terms`` (:122-154, 17 of them
  divides) + 35 in ``geodesic_rhs_scalar`` (:181-197) = 114, plus sin, cos
  and sin(2 theta);
* one Dormand-Prince attempt: 6 right-hand sides + 464 in the stage sums
  (:208-237) + 41 in ``_error_norm`` (:261-268) + 1 in ``_step_factor``
  (:277) = 1190, plus 18 sin/cos and one pow;
* one accepted step: 29 more (affine parameter, Carter and null diagnostics,
  :423-437), plus the 3 trigonometric calls of ``_carter``/``_classify``;
* one photon: 114 for the initial right-hand side and 30 for its Bardeen
  initial condition.

``tests/test_flops.py`` recounts the right-hand-side and stage-sum figures
from the operators of the oracle's restatement and the tableau's nonzero
pattern.
"""

from __future__ import annotations

import numpy as np

RHS_METRIC = 79
RHS_EQUATIONS = 35
RHS = RHS_METRIC + RHS_EQUATIONS
STAGE_SUMS = 464
ERROR_NORM = 41
STEP_FACTOR = 1
PER_ATTEMPT = 6 * RHS + STAGE_SUMS + ERROR_NORM + STEP_FACTOR  # 1190
PER_ACCEPTED = 29
PER_PHOTON = RHS + 30
TRIG_PER_ATTEMPT = 18
POW_PER_ATTEMPT = 1

# Stage sums of the eight-component reference step (per component):
# stage s (1..5) has s terms: 2 s + 1 flops; y_new has 5 terms: 11;
# err has 6 terms times h: 12.
assert STAGE_SUMS == 8 * (sum(2 * s + 1 for s in range(1, 6)) + 11 + 12)
assert PER_ATTEMPT == 1190


def photon_flops(n_steps: np.ndarray, n_rejected: np.ndarray) -> np.ndarray:
    """Algorithmic flops of each photon from its accepted and rejected step counts."""
    n_steps = np.asarray(n_steps, dtype=np.int64)
    n_rejected = np.asarray(n_rejected, dtype=np.int64)
    return PER_PHOTON + PER_ATTEMPT * (n_steps + n_rejected) + PER_ACCEPTED * n_steps


def batch_flops(n_steps: np.ndarray, n_rejected: np.ndarray) -> float:
    return float(np.sum(photon_flops(n_steps, n_rejected), dtype=np.float64))
