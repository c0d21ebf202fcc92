"""Critical impact parameter search (validation/kerr_bisection.py) with a backend argument.

* ``photon_from_constants`` equals the reference's bit for bit on 64
  equatorial inward rays (tests/golden/critical_impacts.npz, written by
  tests/golden/make_golden.py from the reference).
* With the oracle as the backend, the searches reproduce the reference's
  brackets, iteration counts and ray counts, and bracket the closed-form
  b_c of BPT 1972 (eq. 2.18 via ``critical_impact_parameters``).
* Argument checks and the invalid-bracket error of the reference
  (tests/test_kerr.py:121-138).
"""

from __future__ import annotations

import math
import os

import numpy as np
import pytest
from conftest import GOLDEN, options

from kerrray_amd.geodesics import photon_from_constants
from kerrray_amd.geometry import Spacetime
from kerrray_amd.kerr_bisection import bisect_critical_impacts
from kerrray_amd.orbits import critical_impact_parameters

SEARCHES = (
    (0.0, 50.0, 1e-8, 4.0, 7.0, 1e-7, 8, (True,)),
    (0.9, 50.0, 1e-8, 2.0, 8.0, 1e-6, 8, (True, False)),
    (-0.5, 1000.0, 1e-9, 2.0, 9.0, 1e-6, 5, (True, False)),
    (0.998, 50.0, 1e-8, 1.5, 8.0, 1e-5, 1, (True,)),
)


@pytest.fixture(scope="module")
def gold() -> dict:
    with np.load(os.path.join(GOLDEN, "critical_impacts.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_photon_from_constants_equals_the_reference(gold) -> None:
    for (spin, r0, lz), ref in zip(gold["constants_rows"], gold["constants_Y0"]):
        y = photon_from_constants(Spacetime(1.0, spin), r0, 0.5 * math.pi, 1.0, lz, 0.0, sign_r=-1, sign_theta=1)
        assert np.array_equal(y, ref), (spin, r0, lz)


def test_device_probe_states_equal_photon_from_constants(gold) -> None:
    """kr_device.h's impact_state (host build) against photon_from_constants, bit for bit, on the golden rows
    and on a sweep of impact parameters through b_c for several spins."""
    import ctypes
    import subprocess

    from conftest import ROOT

    hostref = os.path.join(ROOT, "tests", "hostref")
    subprocess.run(["make", "-s", "-C", hostref], check=True, capture_output=True)
    L = ctypes.CDLL(os.path.join(hostref, "libkr_hostref.so"))
    L.krh_impact_states.restype = ctypes.c_int
    L.krh_impact_states.argtypes = [ctypes.c_double] * 6 + [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    theta = 0.5 * math.pi
    rows = [(float(s), float(r0), float(lz)) for s, r0, lz in gold["constants_rows"]]
    rng = np.random.default_rng(3)
    for spin in (0.0, 0.5, -0.9, 0.998):
        for r0 in (50.0, 1000.0):
            rows += [(spin, r0, float(v)) for v in rng.uniform(-9.0, 9.0, 64)]
    for spin, r0, lz in rows:
        st = Spacetime(1.0, spin)
        ref = photon_from_constants(st, r0, theta, 1.0, lz, 0.0, sign_r=-1, sign_theta=1)
        got = np.zeros(8)
        arr = np.array([lz])
        assert L.krh_impact_states(st.mass, st.a, r0, theta, math.cos(theta) ** 2, math.sin(theta) ** 2,
                                   arr.ctypes.data, 1, got.ctypes.data) == 0
        assert np.array_equal(got, ref), (spin, r0, lz)


def run(k, backend):
    spin, r0, rtol, lo, hi, tol, n_probe, dirs = SEARCHES[k]
    integ, term = options(rtol=rtol, atol=1e-2 * rtol, escape_radius=r0)
    return Spacetime(1.0, spin), bisect_critical_impacts(Spacetime(1.0, spin), integ, term, r0=r0, b_lo=lo, b_hi=hi,
                                                          tol=tol, max_iterations=60, n_probe=n_probe,
                                                          directions=dirs, backend=backend)


@pytest.mark.parametrize("k", range(len(SEARCHES)))
def test_oracle_backend_reproduces_the_reference_searches(oracle, gold, k) -> None:
    st, res = run(k, oracle)
    ref = gold[f"s{k}_brackets"]
    b_closed = critical_impact_parameters(st)
    for r, row in zip(res, ref):
        assert (r.b_lo, r.b_hi, r.iterations, r.n_rays, r.max_steps) == tuple(row[:2]) + tuple(int(v) for v in row[2:])
        closed = b_closed[0] if r.prograde else b_closed[1]
        assert r.b_lo - 1e-7 <= closed <= r.b_hi + 1e-7


def test_argument_checks(oracle) -> None:
    st = Spacetime(1.0, 0.5)
    integ, term = options(rtol=1e-8, atol=1e-10, escape_radius=50.0)
    with pytest.raises(ValueError):
        bisect_critical_impacts(st, integ, term, r0=50.0, b_lo=5.0, b_hi=4.0, tol=1e-3, max_iterations=5,
                                backend=oracle)
    with pytest.raises(RuntimeError, match="bracket"):
        bisect_critical_impacts(st, integ, term, r0=50.0, b_lo=7.0, b_hi=8.0, tol=1e-3, max_iterations=5,
                                directions=(True,), backend=oracle)
    (res,) = bisect_critical_impacts(st, integ, term, r0=50.0, b_lo=2.0, b_hi=8.0, tol=1e-12, max_iterations=1,
                                     n_probe=5, directions=(True,), backend=oracle)
    assert res.iterations == 1 and res.width <= 6.0 / 6.0 + 1e-12
    assert res.b_lo < critical_impact_parameters(st)[0] < res.b_hi
