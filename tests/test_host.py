"""Host-side logic of the hip backend (no GPU): option validation, constant packing,
initial conditions, registry, chunking and the C ABI library's exported symbols."""

from __future__ import annotations

import ctypes
import math
import os
import re

import numpy as np
import pytest
from conftest import ROOT, ns

from kerrray_amd import _lib
from kerrray_amd.backends import (
    BACKEND_NAMES, Backend, BackendUnavailable, get_backend, kernel_params,
)
from kerrray_amd.camera import BardeenCamera, bardeen_initial_states, bardeen_points
from kerrray_amd.geodesics import EventOptions, IntegratorOptions, TerminationOptions, max_attempts
from kerrray_amd.geometry import Spacetime, outer_horizon
from kerrray_amd.photons import BatchResult
from kerrray_amd.rays import concatenate_results, trace_rays


def test_option_validation() -> None:
    with pytest.raises(ValueError):
        IntegratorOptions(method="euler")
    with pytest.raises(ValueError):
        IntegratorOptions(rtol=0.0)
    with pytest.raises(ValueError):
        IntegratorOptions(h_max=-1.0)
    with pytest.raises(ValueError):
        IntegratorOptions(max_steps=0)
    with pytest.raises(ValueError):
        IntegratorOptions(dtype="float16")
    with pytest.raises(ValueError):
        TerminationOptions(horizon_epsilon=-1.0)
    with pytest.raises(ValueError):
        EventOptions(r_in=5.0, r_out=4.0)
    with pytest.raises(ValueError):
        Spacetime(1.0, 1.0)
    assert max_attempts(IntegratorOptions()) > 1_600_000


def test_kernel_params_match_the_reference_packing(oracle) -> None:
    """kernel_params == numba_backend._kernel_constants + _tableaus (restated in the oracle front end)."""
    for spin, integ, term, ev in [
        (0.9, IntegratorOptions(rtol=1e-6, atol=1e-8), TerminationOptions(), None),
        (-0.7, IntegratorOptions(method="rk4", h_max=0.3, dtype="float32"), TerminationOptions(1e-3, 500.0),
         EventOptions(True, 6.0, 30.0)),
    ]:
        st = Spacetime(1.0, spin)
        p = kernel_params(st, integ, term, ev)
        geom, consts, params, ipar = oracle.kernel_constants(st, integ, term, ev)
        tab_a, tab_e, rk4_b, rk4_c = oracle.tableaus()
        assert np.array_equal(np.frombuffer(p.geom, np.float64), geom)
        assert np.array_equal(np.frombuffer(p.consts, np.float64), consts)
        assert np.array_equal(np.frombuffer(p.params, np.float64), params)
        assert np.array_equal(np.frombuffer(p.ipar, np.int64), ipar)
        assert np.array_equal(np.frombuffer(p.tab_a, np.float64), tab_a.ravel())
        assert np.array_equal(np.frombuffer(p.tab_e, np.float64), tab_e)
        assert np.array_equal(np.frombuffer(p.rk4_b, np.float64), rk4_b)
        assert np.array_equal(np.frombuffer(p.rk4_c, np.float64), rk4_c)
        assert p.dtype == (0 if integ.dtype == "float64" else 1)


def test_bardeen_states_match_the_reference_golden(oracle, golden1, golden2) -> None:
    cam = BardeenCamera(radius=1000.0, inclination_deg=60.0, fov=10.0, resolution=64)
    assert np.array_equal(bardeen_initial_states(cam, Spacetime(1.0, 0.0)), golden1["Y0"])
    cam2 = BardeenCamera(radius=1000.0, inclination_deg=60.0, fov=10.0, resolution=512)
    Y2 = bardeen_initial_states(cam2, Spacetime(1.0, 0.9))
    import hashlib

    assert hashlib.sha256(Y2.tobytes()).hexdigest() == str(golden2["Y0_sha256"])
    assert np.array_equal(Y2, oracle.bardeen_image(1.0, 0.9, 1000.0, 60.0, 10.0, 512))


def test_pixel_grid_is_the_reference_grid() -> None:
    cam = BardeenCamera(radius=1000.0, inclination_deg=60.0, fov=10.0, resolution=8)
    alpha, beta = cam.pixel_coordinates()
    assert alpha.shape == (8, 8) and np.all(alpha[:, ::-1] == -alpha) and np.all(beta[::-1, :] == -beta)
    assert np.all(alpha != 0) and beta[0, 0] > 0 and cam.pixel_size == 2.5
    with pytest.raises(ValueError):
        BardeenCamera(radius=1000.0, inclination_deg=60.0, resolution=7)


def test_bardeen_points_reject_inconsistent_constants() -> None:
    st = Spacetime(1.0, 0.9)
    with pytest.raises(ValueError):
        bardeen_points(st, 1.2, 60.0, [0.5], [0.5])  # inside the horizon (r_+ = 1.436)
    # An image point farther than r_o on the sky has R(r_o) < 0.
    with pytest.raises(ValueError):
        bardeen_points(st, 10.0, 60.0, [40.0], [40.0])


def test_registry() -> None:
    assert BACKEND_NAMES == ("numpy", "numba", "cuda", "hip")
    with pytest.raises(ValueError, match="unknown backend"):
        get_backend("opencl")
    for name in ("numpy", "numba", "cuda"):
        with pytest.raises(BackendUnavailable):
            get_backend(name)
    assert issubclass(BackendUnavailable, RuntimeError)


class _FakeBackend:
    """Host-logic stand-in: returns the input states (exercises chunking only)."""

    name = "fake"

    def integrate_batch(self, st, Y0, integ, term, events=None):
        y = np.asarray(Y0, float)
        n = len(y)
        z = np.zeros(n)
        return BatchResult(Y=y.copy(), state=np.arange(n, dtype=np.int64), n_steps=z.astype(np.int64), lam=z,
                           max_null_error=z, max_energy_drift=z, max_lz_drift=z, max_carter_drift=z,
                           runtime_s=1e-3, n_rejected=z.astype(np.int64))


def test_trace_rays_chunking_and_progress() -> None:
    assert isinstance(_FakeBackend(), Backend)
    Y0 = np.random.default_rng(0).normal(size=(64, 8))
    fractions: list[float] = []
    res = trace_rays(Spacetime(), Y0, IntegratorOptions(), TerminationOptions(), backend=_FakeBackend(),
                     chunk_size=20, progress=fractions.append)
    assert fractions == [20 / 64, 40 / 64, 60 / 64, 1.0]
    assert np.array_equal(res.Y, Y0) and res.n_rays == 64
    with pytest.raises(ValueError):
        trace_rays(Spacetime(), np.zeros((0, 8)), IntegratorOptions(), TerminationOptions(), backend=_FakeBackend())
    with pytest.raises(ValueError):
        concatenate_results([])


class _FakeWriter(_FakeBackend):
    """A backend that writes into the caller's row store (the hip backend's ``integrate_into`` path)."""

    def __init__(self) -> None:
        self.calls: list[int] = []

    def integrate_into(self, st, Y0, integ, term, events, rows) -> None:
        y = np.asarray(Y0, float)
        self.calls.append(len(y))
        assert rows.n == len(y) and rows.Y.base is not None  # a view on the one allocation
        rows.Y[:] = y
        rows.state[:] = np.arange(len(y))


def test_trace_rays_writes_chunks_in_place() -> None:
    Y0 = np.random.default_rng(1).normal(size=(50, 8))
    fake = _FakeWriter()
    res = trace_rays(Spacetime(), Y0, IntegratorOptions(), TerminationOptions(), backend=fake, chunk_size=16)
    assert fake.calls == [16, 16, 16, 2]
    assert np.array_equal(res.Y, Y0)
    assert np.array_equal(res.state, np.concatenate([np.arange(16)] * 3 + [np.arange(2)]))
    assert res.event_Y is None and res.event_hit is None


def test_concatenate_results_keeps_order_and_drops_fields_a_part_lacks() -> None:
    def part(n, start, events, rejected=True):
        z = np.zeros(n)
        return BatchResult(Y=np.full((n, 8), float(start)), state=np.arange(start, start + n, dtype=np.int64),
                           n_steps=np.ones(n, np.int64), lam=z, max_null_error=z, max_energy_drift=z,
                           max_lz_drift=z, max_carter_drift=z, runtime_s=0.5,
                           event_Y=np.zeros((n, 8)) if events else None,
                           event_hit=np.ones(n, bool) if events else None,
                           n_rejected=np.zeros(n, np.int64) if rejected else None)

    a, b = part(3, 0, True), part(2, 3, True)
    joined = concatenate_results([a, b])
    assert np.array_equal(joined.state, np.arange(5)) and joined.runtime_s == 1.0
    assert joined.event_hit.dtype == bool and joined.event_hit.all() and joined.event_Y.shape == (5, 8)
    mixed = concatenate_results([a, part(2, 3, False, rejected=False)], runtime_s=7.0)
    assert mixed.event_Y is None and mixed.event_hit is None and mixed.n_rejected is None
    assert mixed.runtime_s == 7.0 and np.array_equal(mixed.Y[3:, 0], [3.0, 3.0])
    assert concatenate_results([a]) is a


def _header_symbols() -> list[str]:
    text = open(os.path.join(ROOT, "include", "kerrray_hip.h")).read()
    return sorted(set(re.findall(r"^(?:kr_status|int|int64_t|size_t|double|const char\*)\s+(kr_\w+)\(", text, flags=re.M)))


def test_library_loads_and_exports_every_declared_symbol() -> None:
    L = _lib.lib()
    declared = _header_symbols()
    assert set(declared) == set(_lib.EXPORTED)
    for sym in declared:
        assert hasattr(L, sym), sym
    assert L.kr_abi_version() == _lib.ABI_VERSION


def test_null_handle_and_bad_arguments_are_errors_not_crashes() -> None:
    L = _lib.lib()
    p = kernel_params(Spacetime(1.0, 0.9), IntegratorOptions(), TerminationOptions())
    out = _lib.KrOutputs()
    st = L.kr_integrate_batch(None, ctypes.byref(p), None, 0, ctypes.byref(out), None)
    assert st == _lib.KR_ERR_INVALID and "handle" in _lib.last_error()
    assert L.kr_create(0, None) == _lib.KR_ERR_INVALID
    h = ctypes.c_void_p()
    st = L.kr_create(0, ctypes.byref(h))
    if st != _lib.KR_OK:  # no GPU here: the backend is unavailable, not broken
        assert st == _lib.KR_ERR_DEVICE and h.value is None
    else:
        L.kr_destroy(h)


def test_ragged_workload_is_the_committed_table() -> None:
    import json
    from dataclasses import asdict

    from kerrray_amd.batch import ragged_workload

    spec = json.load(open(os.path.join(ROOT, "tests", "golden", "ragged64_spec.json")))
    jobs = ragged_workload(0, 64)
    keys = set(spec["images"][0])
    assert [{k: v for k, v in asdict(j).items() if k in keys} for j in jobs] == spec["images"]
    assert all(j.camera == "bardeen" and j.phi_deg == 0.0 for j in jobs)
    assert sum(j.n_rays for j in jobs) == 5_828_608
    assert {j.inclination_deg for j in jobs} == {17.0, 60.0, 85.0}
    assert {j.resolution for j in jobs} == {64, 128, 256, 512}
    assert all(0.0 <= j.spin < 0.99 and j.atol == 1e-2 * j.rtol for j in jobs)
