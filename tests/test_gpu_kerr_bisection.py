"""Critical impact parameter search on the hip backend (validation/kerr_bisection.py:96-300).

The search runs on the device (kr_impact_search): one probe launch and one
bracket-update launch per round for all searches.

The reference's brackets (tests/golden/critical_impacts.npz) come from its
NumPy integrator. The device classifies every probe as the reference does,
except possibly probes within the integration error of the numerical
threshold: device sin/cos/pow differ from glibc's by an ulp in a few percent
of calls. So the final bracket must lie within one reference bracket width of
the reference's, after the same number of iterations, and contain the closed
form (BPT 1972) to within the integration error.
"""

from __future__ import annotations

import os

import numpy as np
import pytest
from conftest import GOLDEN
from test_kerr_bisection import SEARCHES, run

from kerrray_amd.orbits import critical_impact_parameters

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", range(len(SEARCHES)))
def test_hip_backend_reproduces_the_reference_searches(hip, k) -> None:
    with np.load(os.path.join(GOLDEN, "critical_impacts.npz"), allow_pickle=False) as z:
        ref = z[f"s{k}_brackets"]
    st, res = run(k, hip)
    closed = critical_impact_parameters(st)
    for r, row in zip(res, ref):
        exact = (r.b_lo, r.b_hi) == (row[0], row[1])
        print(f"[critical impact a={st.spin} prograde={r.prograde}] b_c {r.b_c:.10f} (reference "
              f"{0.5 * (row[0] + row[1]):.10f}), bracket identical {exact}, iterations {r.iterations}", flush=True)
        width = row[1] - row[0]
        assert r.iterations == int(row[2]) and r.n_rays == int(row[3])
        assert abs(r.b_lo - row[0]) <= width and abs(r.b_hi - row[1]) <= width
        c = closed[0] if r.prograde else closed[1]
        assert r.b_lo - 1e-7 <= c <= r.b_hi + 1e-7


def test_device_search_errors_and_round_limit(hip) -> None:
    from conftest import options

    from kerrray_amd.geometry import Spacetime
    from kerrray_amd.kerr_bisection import bisect_critical_impacts

    st = Spacetime(1.0, 0.5)
    integ, term = options(rtol=1e-8, atol=1e-10, escape_radius=50.0)
    with pytest.raises(RuntimeError, match="bracket"):
        bisect_critical_impacts(st, integ, term, r0=50.0, b_lo=7.0, b_hi=8.0, tol=1e-3, max_iterations=5,
                                directions=(True,), backend=hip)
    (res,) = bisect_critical_impacts(st, integ, term, r0=50.0, b_lo=2.0, b_hi=8.0, tol=1e-12, max_iterations=1,
                                     n_probe=5, directions=(True,), backend=hip)
    assert res.iterations == 1 and res.width <= 6.0 / 6.0 + 1e-12 and res.n_rays == 2 + 5
    assert res.b_lo < critical_impact_parameters(st)[0] < res.b_hi
